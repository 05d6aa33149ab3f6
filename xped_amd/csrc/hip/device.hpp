// Internal definitions shared by the HIP translation units of the device layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/xped_hip.h"

namespace xh {
// Lifetime of a plan that captured graphs may read: a destroy call on a plan still recorded by a
// live graph is deferred until the last such graph is destroyed (xh_graph_destroy), because the
// graph's kernel nodes hold the plan's device tables by address.
struct PlanLife {
    int pins = 0;       // live graphs (and the capture in progress) that recorded the plan
    bool dead = false;  // destroy was requested while pinned
};
struct PlanPin {
    PlanLife* plan;
    void (*release)(xh_ctx*, PlanLife*);  // frees the plan for good
};
}  // namespace xh

struct xh_ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t events[16] = {};
    // small pinned host staging buffer for scalar reductions
    double* host_scalar = nullptr;
    double* dev_scalar = nullptr;  // RED_PARTS + 64 doubles of device scratch for reductions
    int gemm_tiles = 0;            // FP64 GEMM tiling: 0 128x128 tiles for large sectors, 1 64x64 tiles only
    double* gemm_scratch = nullptr;  // device address of this device's copy of the GEMM scratch array (xh_gemm_noalloc)
    // host staging buffers of asynchronous uploads, kept until the event after their copy fires
    std::vector<std::pair<hipEvent_t, std::vector<char>>> pending_host;
    // pinned host staging buffers for large plan uploads: free ones, and ones whose copy may still run
    struct PinnedBuf {
        void* p = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
    };
    std::vector<PinnedBuf> pinned_free, pinned_busy;
    // pinned slabs that small uploads are sent from: the current one (filled up to slab_used),
    // full ones whose copies may still run, and idle ones
    PinnedBuf slab{};
    size_t slab_used = 0;
    std::vector<PinnedBuf> slab_free, slab_busy;
    // device blocks (xh::dmalloc / xh::dfree): the size class of every block handed out, idle
    // blocks by size class, the bytes they hold and their bound, and counters
    std::unordered_map<void*, size_t> dblock_live;
    std::unordered_map<size_t, std::vector<void*>> dblock_idle;
    size_t dcache_bytes = 0, dcache_limit = size_t(4) << 30;
    int64_t dcache_hits = 0, dcache_misses = 0;
    // HIP-event timing of every grouped GEMM launch (xh_gemm_timing): event pairs in launch order
    bool gemm_timing = false;
    std::vector<hipEvent_t> gemm_events;  // pool, 2 per timed launch
    size_t gemm_events_used = 0;
    std::vector<double> gemm_event_flops;
    // per timed launch: tag, transposes and the (m, n, k) of every item (xh_gemm_launch_log)
    std::vector<std::vector<int32_t>> gemm_launch_items;
    int gemm_tag = 0;  // recorded with each timed launch (xh_set_gemm_tag)
    // the same for permute launches (xh_permute_timing), with the bytes each one moves
    // GEMM tile autotuner: plan signature -> tiling (0 mixed, 1 64x64 only), and counters
    bool gemm_autotune = false;
    std::unordered_map<uint64_t, int> gemm_tune;
    int64_t gemm_tune_runs = 0, gemm_tune_hits = 0;
    // set when a call refuses to run under graph capture (it would allocate or synchronise);
    // xh_graph_end then discards the capture
    bool capture_failed = false;
    std::vector<xh::PlanPin> capture_pins;  // plans recorded by the capture in progress
    bool perm_timing = false;
    std::vector<hipEvent_t> perm_events;
    size_t perm_events_used = 0;
    std::vector<double> perm_event_bytes;
};

// FP64 SVD of large sector blocks by one-sided block Jacobi (bjsvd.hip); same item layout as
// xh_block_svd.
void block_jacobi_svd(xh_ctx* ctx, const std::vector<xh_factor_item>& items, const double* A, double* U, double* S, double* Vt);

namespace xh {

// partial results of one reduction launch (k_maxabs / k_dot) in the context scratch
constexpr int RED_PARTS = 1024;

// take the device address of the current device's copy of the GEMM scratch array (gemm.hip)
void gemm_scratch_init(xh_ctx* ctx);

// while the context stream is being captured: keep the plan alive until the graph is destroyed
void pin_for_capture(xh_ctx* ctx, PlanLife* plan, void (*release)(xh_ctx*, PlanLife*));
void unpin(xh_ctx* ctx, const PlanPin& p);

// refuse work that cannot be recorded into a graph while the context stream is being captured
void refuse_under_capture(xh_ctx* ctx, const char* what);

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what)
{
    if(e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? XH_ENOMEM : XH_EHIP;
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Runs f and converts exceptions into status codes stored on the context.
template <typename F>
int guard(xh_ctx* ctx, F&& f)
{
    try {
        f();
        return XH_OK;
    } catch(const Error& e) {
        if(ctx) ctx->err = e.what();
        return e.code;
    } catch(const std::bad_alloc& e) {
        if(ctx) ctx->err = "host allocation failed";
        return XH_ENOMEM;
    } catch(const std::exception& e) {
        if(ctx) ctx->err = e.what();
        return XH_EINVAL;
    }
}

inline void launch_check(const char* name) { check(hipGetLastError(), name); }

// event pair around one timed GEMM launch on the context stream
inline void gemm_timer_begin(xh_ctx* ctx)
{
    while(ctx->gemm_events.size() < ctx->gemm_events_used + 2) {
        hipEvent_t e;
        check(hipEventCreate(&e), "hipEventCreate");
        ctx->gemm_events.push_back(e);
    }
    check(hipEventRecord(ctx->gemm_events[ctx->gemm_events_used], ctx->stream), "hipEventRecord");
}
inline void gemm_timer_end(xh_ctx* ctx, double flops, std::vector<int32_t>&& items = {})
{
    ctx->gemm_launch_items.push_back(std::move(items));
    check(hipEventRecord(ctx->gemm_events[ctx->gemm_events_used + 1], ctx->stream), "hipEventRecord");
    ctx->gemm_events_used += 2;
    ctx->gemm_event_flops.push_back(flops);
}
inline void perm_timer_begin(xh_ctx* ctx)
{
    while(ctx->perm_events.size() < ctx->perm_events_used + 2) {
        hipEvent_t e;
        check(hipEventCreate(&e), "hipEventCreate");
        ctx->perm_events.push_back(e);
    }
    check(hipEventRecord(ctx->perm_events[ctx->perm_events_used], ctx->stream), "hipEventRecord");
}
inline void perm_timer_end(xh_ctx* ctx, double bytes)
{
    check(hipEventRecord(ctx->perm_events[ctx->perm_events_used + 1], ctx->stream), "hipEventRecord");
    ctx->perm_events_used += 2;
    ctx->perm_event_bytes.push_back(bytes);
}

// Round-up magic-number division for 0 <= x < 2^31 and 1 <= d < 2^31
// (Granlund & Montgomery, "Division by invariant integers using multiplication", PLDI 1994).
struct FastDiv {
    uint32_t d, magic, shift;
};
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{d, 0, 0};
    uint32_t l = 0;
    while((uint64_t(1) << l) < d) ++l;
    f.shift = l;
    f.magic = static_cast<uint32_t>(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t x, uint32_t magic, uint32_t shift)
{
    return (__umulhi(x, magic) + x) >> shift;
}

// One output leg of a permute task (shared by the permute kernels and the GEMM that reads its A
// operand through a permute): extent, magic division by the extent, and the stride of the leg's
// source leg in the source subblock's row (bit 8 of flags clear) or column (set) index; bits 0-7
// of flags: division shift.
struct LegInfo {
    uint32_t ext, magic, flags, stride;
};
__host__ __device__ __forceinline__ uint32_t leg_shift(const LegInfo& l) { return l.flags & 255u; }
__host__ __device__ __forceinline__ int64_t leg_rstride(const LegInfo& l) { return (l.flags & 256u) ? 0 : int64_t(l.stride); }
__host__ __device__ __forceinline__ int64_t leg_cstride(const LegInfo& l) { return (l.flags & 256u) ? int64_t(l.stride) : 0; }
inline LegInfo make_leg(uint32_t ext, int64_t sstride, bool is_col)
{
    const FastDiv f = make_fastdiv(ext);
    return LegInfo{ext, f.magic, f.shift | (is_col ? 256u : 0u), static_cast<uint32_t>(sstride)};
}

// Stream-ordered device memory of the context (ctx.hip).  hipFreeAsync took 90 us of host time
// per call on the MI355X (212 ms in one CTMRG step of 2361 frees), so freed blocks are kept in
// size classes (four per power of two) and handed out again by later requests of the same class.
// Every launch of a context runs on its one stream, so a block freed after the launches that read
// it were enqueued may serve the next launch at once, as with hipFreeAsync.  Idle blocks beyond
// dcache_limit bytes go back with hipFreeAsync; blocks requested under graph capture bypass the
// cache.
hipError_t dmalloc(xh_ctx* ctx, void** p, size_t bytes);
hipError_t dfree(xh_ctx* ctx, void* p);
void dcache_release(xh_ctx* ctx);  // every idle block back to the runtime

inline void dev_free(xh_ctx* ctx, void* p)
{
    if(p) dfree(ctx, p);
}

// Pinned staging buffers (hipHostMalloc), reused once the copy that read them has completed.
inline void pinned_reclaim(xh_ctx* ctx)
{
    auto& busy = ctx->pinned_busy;
    std::size_t keep = 0;
    for(std::size_t i = 0; i < busy.size(); ++i) {
        if(hipEventQuery(busy[i].ev) == hipSuccess)
            ctx->pinned_free.push_back(busy[i]);
        else
            busy[keep++] = busy[i];
    }
    busy.resize(keep);
}
inline xh_ctx::PinnedBuf pinned_get(xh_ctx* ctx, size_t bytes)
{
    pinned_reclaim(ctx);
    size_t cap = size_t(1) << 20;
    while(cap < bytes) cap <<= 1;
    auto& fr = ctx->pinned_free;
    for(std::size_t i = 0; i < fr.size(); ++i)
        if(fr[i].cap == cap) {
            xh_ctx::PinnedBuf b = fr[i];
            fr[i] = fr.back();
            fr.pop_back();
            return b;
        }
    // keep at most 16 idle buffers
    while(fr.size() >= 16) {
        hipHostFree(fr.back().p);
        hipEventDestroy(fr.back().ev);
        fr.pop_back();
    }
    xh_ctx::PinnedBuf b;
    b.cap = cap;
    check(hipHostMalloc(&b.p, cap, hipHostMallocDefault), "hipHostMalloc(staging)");
    check(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming), "hipEventCreate");
    return b;
}
// the buffer becomes reusable after everything enqueued so far on the stream (its copy) is done
inline void pinned_put(xh_ctx* ctx, const xh_ctx::PinnedBuf& b)
{
    check(hipEventRecord(b.ev, ctx->stream), "hipEventRecord");
    ctx->pinned_busy.push_back(b);
}
inline void pinned_destroy_all(xh_ctx* ctx)
{
    if(ctx->slab.p) ctx->slab_free.push_back(ctx->slab);
    ctx->slab = {};
    for(auto* v : {&ctx->pinned_free, &ctx->pinned_busy, &ctx->slab_free, &ctx->slab_busy})
        for(auto& b : *v) {
            hipHostFree(b.p);
            hipEventDestroy(b.ev);
        }
    ctx->pinned_free.clear();
    ctx->pinned_busy.clear();
    ctx->slab_free.clear();
    ctx->slab_busy.clear();
}

// A host-to-device copy from pageable memory may return only once the work queued before it on
// the stream has run (a 200 KB copy behind 30 ms of queued kernels returned after 8 ms on the
// MI355X, the same copy from pinned memory after 0.01 ms), which starves the GPU while plans are
// built between launches.  Small uploads are therefore copied into a pinned slab and sent from
// there: each takes a 256-byte aligned region of the current slab; a full slab is retired behind
// an event recorded after its last copy and reused once that event has completed.
constexpr size_t SLAB_BYTES = size_t(4) << 20;
inline void* slab_take(xh_ctx* ctx, size_t bytes)
{
    if(bytes > SLAB_BYTES) throw Error(XH_EINVAL, "slab upload too large");
    size_t off = (ctx->slab_used + 255) & ~size_t(255);
    if(!ctx->slab.p || off + bytes > ctx->slab.cap) {
        if(ctx->slab.p) {
            check(hipEventRecord(ctx->slab.ev, ctx->stream), "hipEventRecord");
            ctx->slab_busy.push_back(ctx->slab);
            ctx->slab = {};
        }
        auto& busy = ctx->slab_busy;
        std::size_t keep = 0;
        for(std::size_t i = 0; i < busy.size(); ++i) {
            if(hipEventQuery(busy[i].ev) == hipSuccess)
                ctx->slab_free.push_back(busy[i]);
            else
                busy[keep++] = busy[i];
        }
        busy.resize(keep);
        while(ctx->slab_free.size() > 8) {  // keep at most 8 idle slabs
            hipHostFree(ctx->slab_free.back().p);
            hipEventDestroy(ctx->slab_free.back().ev);
            ctx->slab_free.pop_back();
        }
        if(!ctx->slab_free.empty()) {
            ctx->slab = ctx->slab_free.back();
            ctx->slab_free.pop_back();
        } else {
            xh_ctx::PinnedBuf b;
            b.cap = SLAB_BYTES;
            check(hipHostMalloc(&b.p, b.cap, hipHostMallocDefault), "hipHostMalloc(slab)");
            if(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
                hipHostFree(b.p);
                throw Error(XH_EHIP, "hipEventCreate(slab)");
            }
            ctx->slab = b;
        }
        off = 0;
    }
    ctx->slab_used = off + bytes;
    return static_cast<char*>(ctx->slab.p) + off;
}

// Asynchronous copy of host memory to the device, ordered on the context stream; the caller may
// reuse or free src as soon as this returns.  Not for a stream under graph capture (the graph
// would copy from the staging memory again at every replay).
inline void upload_async(xh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if(bytes == 0) return;
    if(bytes <= SLAB_BYTES / 4) {
        void* s = slab_take(ctx, bytes);
        std::memcpy(s, src, bytes);
        check(hipMemcpyAsync(dst, s, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync(upload)");
        return;
    }
    xh_ctx::PinnedBuf b = pinned_get(ctx, bytes);
    std::memcpy(b.p, src, bytes);
    const hipError_t e = hipMemcpyAsync(dst, b.p, bytes, hipMemcpyHostToDevice, ctx->stream);
    pinned_put(ctx, b);
    check(e, "hipMemcpyAsync(upload)");
}

// a device copy of v, freed by the caller (dev_free) in stream order
template <typename T>
T* dev_upload(xh_ctx* ctx, const std::vector<T>& v)
{
    if(v.empty()) return nullptr;
    refuse_under_capture(ctx, "a host upload");
    T* p = nullptr;
    check(xh::dmalloc(ctx, reinterpret_cast<void**>(&p), v.size() * sizeof(T)), "hipMallocAsync(plan)");
    upload_async(ctx, p, v.data(), v.size() * sizeof(T));
    return p;
}

// Host arrays packed into one device allocation and sent with one asynchronous copy on the
// context stream, without a host synchronisation.  The host arrays are sent through upload_async
// (pinned staging) and dropped; only under graph capture is the host copy kept until release(),
// which first waits for the event recorded after the copy.  Large uploads whose size is known up
// front (begin_pinned) are written straight into a pinned staging buffer that returns to the
// context's pool when the copy has completed.
struct PackedUpload {
    std::vector<char> host;
    void* dev = nullptr;
    hipEvent_t done = nullptr;
    xh_ctx* owner = nullptr;
    xh_ctx::PinnedBuf pin{};
    size_t pin_used = 0;
    PackedUpload() = default;
    PackedUpload(const PackedUpload&) = delete;
    PackedUpload& operator=(const PackedUpload&) = delete;
    ~PackedUpload()
    {
        if(owner && pin.p) pinned_put(owner, pin);  // error path before the upload
        pin.p = nullptr;
        if(owner && (dev || done)) release(owner);  // error paths: wait and free
    }
    void begin_pinned(xh_ctx* ctx, size_t bytes)
    {
        owner = ctx;
        pin = pinned_get(ctx, bytes);
        pin_used = 0;
    }
    template <typename T>
    size_t add(const std::vector<T>& v)
    {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + v.size() * sizeof(T));
        if(!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    // an aligned region of n elements of T, filled through host_ptr before upload
    template <typename T>
    size_t reserve(int64_t n)
    {
        const size_t bytes = size_t(std::max<int64_t>(n, 0)) * sizeof(T);
        if(pin.p) {
            const size_t off = (pin_used + 255) & ~size_t(255);
            if(off + bytes > pin.cap) throw Error(XH_EINVAL, "staging buffer too small");
            pin_used = off + bytes;
            return off;
        }
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + bytes);
        return off;
    }
    // bytes that reserve() regions of the given sizes take, with their alignment
    static size_t packed_bytes(std::initializer_list<size_t> sizes)
    {
        size_t t = 0;
        for(size_t b : sizes) t = ((t + 255) & ~size_t(255)) + b;
        return t;
    }
    template <typename T>
    T* host_ptr(size_t off)
    {
        return reinterpret_cast<T*>((pin.p ? static_cast<char*>(pin.p) : host.data()) + off);
    }
    void upload(xh_ctx* ctx)
    {
        if(pin.p) {
            owner = ctx;
            if(pin_used > 0) {
                check(xh::dmalloc(ctx, &dev, pin_used), "hipMallocAsync(plan)");
                check(hipMemcpyAsync(dev, pin.p, pin_used, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync(plan)");
            }
            pinned_put(ctx, pin);
            pin.p = nullptr;
            return;
        }
        if(host.empty()) return;
        owner = ctx;
        check(xh::dmalloc(ctx, &dev, host.size()), "hipMallocAsync(plan)");
        if(!xh_graph_capturing(ctx)) {
            upload_async(ctx, dev, host.data(), host.size());
            std::vector<char>().swap(host);
            return;
        }
        // under capture the graph's copy node reads the host copy, kept until release()
        check(hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync(plan)");
        check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreate");
        check(hipEventRecord(done, ctx->stream), "hipEventRecord");
    }
    template <typename T>
    T* ptr(size_t off) const
    {
        return dev ? reinterpret_cast<T*>(static_cast<char*>(dev) + off) : nullptr;
    }
    void release(xh_ctx* ctx)
    {
        if(done) {
            hipEventSynchronize(done);
            hipEventDestroy(done);
            done = nullptr;
        }
        if(dev) xh::dfree(ctx, dev);
        dev = nullptr;
        host.clear();
    }
    // after the kernels that read the device copy have been enqueued: the device memory is
    // freed in stream order and the host copy handed to the context until its copy has finished
    void release_deferred(xh_ctx* ctx);
};

// drop the staging buffers whose copies have completed (wait = true: all of them, after a sync)
inline void collect_staging(xh_ctx* ctx, bool wait)
{
    auto& v = ctx->pending_host;
    std::size_t keep = 0;
    for(std::size_t i = 0; i < v.size(); ++i) {
        bool ready = true;
        if(wait)
            hipEventSynchronize(v[i].first);
        else
            ready = hipEventQuery(v[i].first) == hipSuccess;
        if(ready) {
            hipEventDestroy(v[i].first);
        } else {
            if(keep != i) v[keep] = std::move(v[i]);
            ++keep;
        }
    }
    v.resize(keep);
}

inline void PackedUpload::release_deferred(xh_ctx* ctx)
{
    if(dev) xh::dfree(ctx, dev);
    dev = nullptr;
    if(done) ctx->pending_host.emplace_back(done, std::move(host));
    done = nullptr;
    host.clear();
    if(ctx->pending_host.size() > 64) collect_staging(ctx, false);
}

}  // namespace xh
