// Device-resident tensor operations: structure on the host, data movement and arithmetic on the GPU.
#include <cstdlib>
#include "tensor.hpp"

#include "ad.hpp"

#include "../common/pool.hpp"

#include <cstddef>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>
#include <unordered_map>
#include <stdexcept>

namespace xa {

namespace {

void dcheck(xh_ctx* ctx, int rc, const char* what)
{
    if(rc != XH_OK) throw Error(rc, std::string(what) + ": " + xh_last_error(ctx));
}

void put_i(std::string& s, long long v) { s.append(reinterpret_cast<const char*>(&v), sizeof(v)); }

void put_basis(std::string& s, const Qbasis& b)
{
    put_i(s, b.conj_flag ? 1 : 0);
    put_i(s, static_cast<long long>(b.data.size()));
    for(const auto& e : b.data) {
        put_i(s, e.q.v[0]);
        put_i(s, e.q.v[1]);
        put_i(s, static_cast<long long>(e.dim));
        // a depth-1 tree records the duality of the leg
        put_i(s, e.trees.empty() ? 0 : (e.trees[0].dual[0] ? 1 : 0));
    }
}

// the charges (not the dimensions) of every leg: equal keys give equal tree enumerations
void put_basis_charges(std::string& s, const Qbasis& b)
{
    put_i(s, b.conj_flag ? 1 : 0);
    put_i(s, static_cast<long long>(b.data.size()));
    for(const auto& e : b.data) {
        put_i(s, e.q.v[0]);
        put_i(s, e.q.v[1]);
        put_i(s, e.trees.empty() ? 0 : (e.trees[0].dual[0] ? 1 : 0));
    }
}

std::string charges_key(const TensorStructure& t)
{
    std::string k;
    put_i(k, t.sym.kind[0]);
    put_i(k, t.sym.kind[1]);
    put_i(k, static_cast<long long>(t.unc_dom.size()));
    put_i(k, static_cast<long long>(t.unc_cod.size()));
    for(const auto& b : t.unc_dom) put_basis_charges(k, b);
    for(const auto& b : t.unc_cod) put_basis_charges(k, b);
    return k;
}

}  // namespace

Runtime::~Runtime()
{
    if(ctx) {
        for(auto& [k, v] : gather_cache) v->release(ctx);  // before the GEMM plans they refer to
        for(auto& [k, v] : permute_cache)
            if(v->dev) xh_permute_plan_destroy(ctx, v->dev);
        for(auto& [k, v] : gemm_cache)
            if(v->dev) xh_gemm_plan_destroy(ctx, v->dev);
    }
}

namespace {
// drop the least recently used half of a plan cache; entries still held elsewhere are kept
template <class Map, class Destroy>
long trim_lru(Map& m, std::size_t max_entries, Destroy destroy)
{
    if(m.size() <= max_entries) return 0;
    std::vector<std::pair<std::uint64_t, typename Map::iterator>> byuse;
    byuse.reserve(m.size());
    for(auto it = m.begin(); it != m.end(); ++it) byuse.emplace_back(it->second->last_use, it);
    std::sort(byuse.begin(), byuse.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    long n = 0;
    const std::size_t drop = m.size() - max_entries / 2;
    for(std::size_t i = 0; i < drop && i < byuse.size(); ++i) {
        auto it = byuse[i].second;
        if(it->second.use_count() > 1) continue;
        destroy(*it->second);
        m.erase(it);
        ++n;
    }
    return n;
}
}  // namespace

void Runtime::trim_caches()
{
    n_evicted += trim_lru(gather_cache, max_plans, [this](CachedGather& c) {
        if(ctx) c.release(ctx);
    });
    n_evicted += trim_lru(permute_cache, max_plans, [this](CachedPermute& c) {
        if(ctx && c.dev) xh_permute_plan_destroy(ctx, c.dev);
        c.dev = nullptr;
    });
    n_evicted += trim_lru(gemm_cache, max_plans, [this](CachedGemm& c) {
        if(ctx && c.dev) xh_gemm_plan_destroy(ctx, c.dev);
        c.dev = nullptr;
    });
    // host-only caches: dimension-free permute parts and shared structures
    if(sym_permute_cache.size() > 2 * max_plans) sym_permute_cache.clear();
    if(struct_cache.size() > 4 * max_plans) struct_cache.clear();
}

void CachedGather::release(xh_ctx* ctx)
{
    for(auto& [k, sc] : scatter)
        if(sc.dev) xh_scatter_plan_destroy(ctx, sc.dev);
    scatter.clear();
    if(dev) xh_gather_plan_destroy(ctx, dev);
    dev = nullptr;
}

void Runtime::require_device() const
{
    if(!ctx) throw Error(XH_ENODEV, "this runtime has no GPU (structure-only)");
}

std::string legs_key(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c)
{
    std::string k;
    put_i(k, s.kind[0]);
    put_i(k, s.kind[1]);
    put_i(k, static_cast<long long>(d.size()));
    put_i(k, static_cast<long long>(c.size()));
    for(const auto& b : d) put_basis(k, b);
    for(const auto& b : c) put_basis(k, b);
    return k;
}

std::string structure_key(const TensorStructure& t)
{
    std::string k = legs_key(t.sym, t.unc_dom, t.unc_cod);
    put_i(k, static_cast<long long>(t.sector.size()));
    for(const Q& q : t.sector) {
        put_i(k, q.v[0]);
        put_i(k, q.v[1]);
    }
    return k;
}

StructPtr cached_structure(Runtime* rt, const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c)
{
    std::string k = legs_key(s, d, c);
    auto it = rt->struct_cache.find(k);
    if(it != rt->struct_cache.end()) return it->second;
    StructPtr st = TensorStructure::create(s, d, c);
    rt->struct_cache.emplace(k, st);
    return st;
}

// ------------------------------------------------------------------ DTensor basics

DTensor DTensor::empty(Runtime* rt, StructPtr st)
{
    DTensor t;
    t.rt = rt;
    t.st = std::move(st);
    t.buf = std::make_shared<DeviceBuffer>();
    if(rt->ctx && t.st->size > 0) {
        t.buf->holder = rt->holder;
        t.buf->n = t.st->size;
        void* p = nullptr;
        dcheck(rt->ctx, xh_alloc(rt->ctx, t.st->size * sizeof(double), &p), "allocation of tensor data");
        t.buf->ptr = static_cast<double*>(p);
    }
    return t;
}

DTensor DTensor::zeros(Runtime* rt, StructPtr st)
{
    DTensor t = empty(rt, std::move(st));
    if(rt->ctx && t.size() > 0) dcheck(rt->ctx, xh_fill(rt->ctx, t.data(), t.size(), 0.0), "setZero");
    return t;
}

void DTensor::upload(const double* host, std::size_t n)
{
    rt->require_device();
    if(n != size()) throw Error(XH_EINVAL, "upload: size mismatch");
    dcheck(rt->ctx, xh_upload(rt->ctx, data(), host, n * sizeof(double)), "upload");
}

void DTensor::download(double* host, std::size_t n) const
{
    rt->require_device();
    if(n != size()) throw Error(XH_EINVAL, "download: size mismatch");
    dcheck(rt->ctx, xh_download(rt->ctx, host, data(), n * sizeof(double)), "download");
}

void DTensor::set_random(std::mt19937& engine, int dist)
{
    std::vector<double> h(size());
    if(dist == 0) {
        std::uniform_real_distribution<double> d(-1., 1.);
        for(auto& x : h) x = d(engine);
    } else {
        std::normal_distribution<double> d(0., 1.);
        for(auto& x : h) x = d(engine);
    }
    upload(h.data(), h.size());
}

// ------------------------------------------------------------------ permute

template <class Lap>
static void finish_permute_plan(Runtime* rt, CachedPermute& cpr, const TensorStructure& in, Lap& lap);

// The dimension-free part of a permute plan, shared by all structures with the same charges
// (tensors whose sector dimensions changed after a truncation); only offsets and extents are
// recomputed per structure.
static std::shared_ptr<const SymbolicPermute> symbolic_permute(Runtime* rt, const TensorStructure& in, const TensorStructure& out, int shift,
                                                               const std::vector<int>& p)
{
    std::string sk = charges_key(in);
    put_i(sk, shift);
    for(int x : p) put_i(sk, x);
    auto si = rt->sym_permute_cache.find(sk);
    if(si != rt->sym_permute_cache.end()) {
        rt->n_symbolic_hits++;
        return si->second;
    }
    auto sp = std::make_shared<const SymbolicPermute>(make_symbolic_permute(in, out, shift, p));
    rt->sym_permute_cache.emplace(sk, sp);
    return sp;
}

std::shared_ptr<CachedPermute> get_permute_plan(Runtime* rt, const TensorStructure& in, int shift, const std::vector<int>& p)
{
    std::string k = structure_key(in);
    put_i(k, shift);
    for(int x : p) put_i(k, x);
    auto it = rt->permute_cache.find(k);
    if(it != rt->permute_cache.end()) {
        it->second->last_use = ++rt->tick;
        return it->second;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto lap = [rt, t = t0](int slot) mutable {
        const auto now = std::chrono::steady_clock::now();
        rt->host_t[slot] += std::chrono::duration<double>(now - t).count();
        t = now;
    };
    auto cp = std::make_shared<CachedPermute>();
    cp->last_use = ++rt->tick;
    StructPtr out = permuted_structure(in, shift, p);
    lap(5);
    std::shared_ptr<const SymbolicPermute> sp = symbolic_permute(rt, in, *out, shift, p);
    lap(6);
    cp->plan = numeric_permute_plan(*sp, in, out, p);
    lap(0);
    finish_permute_plan(rt, *cp, in, lap);
    rt->permute_cache.emplace(k, cp);
    rt->trim_caches();
    rt->permute_plan_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rt->n_permute_plans++;
    return cp;
}

// Bounds checks, output-structure sharing and the device upload of a host permute plan.
template <class Lap>
static void finish_permute_plan(Runtime* rt, CachedPermute& cpr, const TensorStructure& in, Lap& lap)
{
    CachedPermute* cp = &cpr;
    // every access of the plan must stay inside the input and output buffers
    {
        const PermutePlan& pl = cp->plan;
        const std::int64_t in_size = static_cast<std::int64_t>(in.size), out_size = static_cast<std::int64_t>(pl.out->size);
        xp::parallel_ranges(pl.tasks.size(), 8192, [&](std::size_t t0, std::size_t t1) {
            for(std::size_t t = t0; t < t1; ++t) {
                const PermuteTask& tk = pl.tasks[t];
                if(tk.rows <= 0 || tk.cols <= 0) throw Error(XH_EINVAL, "permute plan: empty subblock");
                const std::int64_t dmax = tk.dst_off + (tk.rows - 1) + (tk.cols - 1) * tk.dst_ld;
                if(tk.dst_off < 0 || dmax >= out_size) throw Error(XH_EINVAL, "permute plan: output subblock out of bounds");
                std::int64_t rmax = 0, cmax = 0;
                for(int k2 = 0; k2 < pl.nlegs; ++k2) (tk.is_col[k2] ? cmax : rmax) += (static_cast<std::int64_t>(tk.ext[k2]) - 1) * tk.sstride[k2];
                if(tk.src_begin < 0 || tk.nsrc < 0 || static_cast<std::size_t>(tk.src_begin) + static_cast<std::size_t>(tk.nsrc) > pl.sources.size())
                    throw Error(XH_EINVAL, "permute plan: source range out of bounds");
                for(int s2 = 0; s2 < tk.nsrc; ++s2) {
                    const PermuteSource& so = pl.sources[tk.src_begin + s2];
                    if(so.src_off < 0 || so.src_off + rmax + cmax * so.src_ld >= in_size || rmax >= so.src_ld)
                        throw Error(XH_EINVAL, "permute plan: source subblock out of bounds");
                }
            }
        });
    }
    // share output structures between identical leg sets
    {
        std::string sk = legs_key(cp->plan.out->sym, cp->plan.out->unc_dom, cp->plan.out->unc_cod);
        auto si = rt->struct_cache.find(sk);
        if(si == rt->struct_cache.end())
            rt->struct_cache.emplace(sk, cp->plan.out);
        else if(si->second->sector == cp->plan.out->sector)
            cp->plan.out = si->second;
    }
    if(rt->ctx) {
        // the host plan's task and source records have the layout of the C ABI's descriptors
        static_assert(sizeof(PermuteTask) == sizeof(xh_permute_task) && offsetof(PermuteTask, ext) == offsetof(xh_permute_task, ext) &&
                          offsetof(PermuteTask, sstride) == offsetof(xh_permute_task, sstride) &&
                          offsetof(PermuteTask, is_col) == offsetof(xh_permute_task, is_col) && offsetof(PermuteTask, nsrc) == offsetof(xh_permute_task, nsrc),
                      "PermuteTask must match xh_permute_task");
        static_assert(sizeof(PermuteSource) == sizeof(xh_permute_source) && offsetof(PermuteSource, coeff) == offsetof(xh_permute_source, coeff),
                      "PermuteSource must match xh_permute_source");
        xh_permute_desc desc{cp->plan.nlegs, cp->plan.new_rank, static_cast<int64_t>(cp->plan.tasks.size()), static_cast<int64_t>(cp->plan.sources.size()),
                             reinterpret_cast<xh_permute_task*>(cp->plan.tasks.data()), reinterpret_cast<xh_permute_source*>(cp->plan.sources.data())};
        lap(1);
        dcheck(rt->ctx, xh_permute_plan_create(rt->ctx, &desc, &cp->dev), "permute plan");
        recycle_plan_tables(cp->plan);  // the device plan holds its own copy
        lap(2);
    }
}

std::shared_ptr<CachedPermute> get_permute_adj_plan(Runtime* rt, const TensorStructure& y, int shift, const std::vector<int>& p)
{
    std::string k = "adj|" + structure_key(y);
    put_i(k, shift);
    for(int x : p) put_i(k, x);
    auto it = rt->permute_cache.find(k);
    if(it != rt->permute_cache.end()) {
        it->second->last_use = ++rt->tick;
        return it->second;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto lap = [rt, t = t0](int slot) mutable {
        const auto now = std::chrono::steady_clock::now();
        rt->host_t[slot] += std::chrono::duration<double>(now - t).count();
        t = now;
    };
    auto cp = std::make_shared<CachedPermute>();
    cp->last_use = ++rt->tick;
    StructPtr xs = permute_adj_structure(y, shift, p);
    cp->plan = make_permute_adj_plan(xs, y, shift, p);
    lap(0);
    finish_permute_plan(rt, *cp, y, lap);
    rt->permute_cache.emplace(k, cp);
    rt->trim_caches();
    rt->permute_plan_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rt->n_permute_plans++;
    return cp;
}

DTensor permute_adj(const DTensor& y, int shift, const std::vector<int>& p)
{
    // Tensor.cpp:264: the identity returns the tensor itself
    bool ident = (shift == 0);
    for(std::size_t i = 0; ident && i < p.size(); ++i) ident = (p[i] == static_cast<int>(i));
    if(ident && static_cast<int>(p.size()) == y.st->rank + y.st->corank) return y;
    auto cp = get_permute_adj_plan(y.rt, *y.st, shift, p);
    DTensor out = DTensor::empty(y.rt, cp->plan.out);
    if(y.rt->ctx && out.size() > 0) {
        dcheck(y.rt->ctx, xh_permute(y.rt->ctx, cp->dev, y.data(), out.data()), "permute_adj");
        double b = 0;
        xh_permute_plan_bytes(cp->dev, &b);
        y.rt->permute_bytes += b;
        y.rt->n_permute++;
    }
    return out;
}

DTensor permute(const DTensor& t, int shift, const std::vector<int>& p)
{
    // identity permutation without shift returns the tensor itself (Tensor.cpp:344)
    bool ident = (shift == 0);
    for(std::size_t i = 0; ident && i < p.size(); ++i) ident = (p[i] == static_cast<int>(i));
    if(ident && static_cast<int>(p.size()) == t.st->rank + t.st->corank) return t;
    auto cp = get_permute_plan(t.rt, *t.st, shift, p);
    DTensor out = DTensor::empty(t.rt, cp->plan.out);
    if(t.rt->ctx && out.size() > 0) {
        dcheck(t.rt->ctx, xh_permute(t.rt->ctx, cp->dev, t.data(), out.data()), "permute");
        double b = 0;
        xh_permute_plan_bytes(cp->dev, &b);
        t.rt->permute_bytes += b;
        t.rt->n_permute++;
    }
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out, shift, p](Tape& tape) {
            // ADTensor.hpp:171-185: curr.adj() += res.adj().permute_adj<shift, p...>()
            if(const DTensor* d = tape.find(out)) tape.accumulate(t, permute_adj(*d, shift, p));
        });
    return out;
}

// ------------------------------------------------------------------ operator* and contract

static std::shared_ptr<CachedGemm> get_gemm_plan(Runtime* rt, const TensorStructure& l, const TensorStructure& r)
{
    std::string k = structure_key(l);
    k += "|";
    k += structure_key(r);
    auto it = rt->gemm_cache.find(k);
    if(it != rt->gemm_cache.end()) {
        it->second->last_use = ++rt->tick;
        return it->second;
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto lap = [rt, t = t0](int slot) mutable {
        const auto now = std::chrono::steady_clock::now();
        rt->host_t[slot] += std::chrono::duration<double>(now - t).count();
        t = now;
    };
    auto cg = std::make_shared<CachedGemm>();
    cg->last_use = ++rt->tick;
    cg->plan = make_gemm_plan(l, r);
    for(const auto& g : cg->plan.items) {
        const std::int64_t mk = static_cast<std::int64_t>(g.m) * g.k, kn = static_cast<std::int64_t>(g.k) * g.n, mn = static_cast<std::int64_t>(g.m) * g.n;
        if(g.a_off < 0 || g.a_off + mk > static_cast<std::int64_t>(l.size) || g.b_off < 0 || g.b_off + kn > static_cast<std::int64_t>(r.size) || g.c_off < 0 ||
           g.c_off + mn > static_cast<std::int64_t>(cg->plan.out->size))
            throw Error(XH_EINVAL, "gemm plan: block out of bounds");
    }
    {
        std::string sk = legs_key(cg->plan.out->sym, cg->plan.out->unc_dom, cg->plan.out->unc_cod);
        auto si = rt->struct_cache.find(sk);
        if(si == rt->struct_cache.end())
            rt->struct_cache.emplace(sk, cg->plan.out);
        else
            cg->plan.out = si->second;
    }
    const TensorStructure& o = *cg->plan.out;
    std::vector<char> written(o.sector.size(), 0);
    for(const auto& g : cg->plan.items) {
        for(std::size_t s = 0; s < o.sector.size(); ++s)
            if(static_cast<std::int64_t>(o.offsets[s]) == g.c_off) written[s] = 1;
    }
    for(std::size_t s = 0; s < o.sector.size(); ++s)
        if(!written[s] && o.rows[s] * o.cols[s] > 0) {
            xh_copy_rect r{0, 0, static_cast<int64_t>(o.offsets[s]), static_cast<int64_t>(o.rows[s]), static_cast<int32_t>(o.rows[s]),
                           static_cast<int32_t>(o.cols[s]), 0.0, 0};
            cg->zero_rects.push_back(r);
            const std::int64_t off = static_cast<std::int64_t>(o.offsets[s]), len = static_cast<std::int64_t>(o.rows[s] * o.cols[s]);
            auto& zr = cg->zero_ranges;
            if(!zr.empty() && zr[zr.size() - 2] + zr.back() == off)
                zr.back() += len;
            else {
                zr.push_back(off);
                zr.push_back(len);
            }
        }
    lap(3);
    if(rt->ctx) {
        std::vector<xh_gemm_item> items;
        for(const auto& g : cg->plan.items) items.push_back(xh_gemm_item{g.a_off, g.b_off, g.c_off, g.m, g.n, g.k});
        dcheck(rt->ctx, xh_gemm_plan_create(rt->ctx, items.data(), static_cast<int64_t>(items.size()), 0, 0, &cg->dev), "gemm plan");
    }
    lap(4);
    rt->gemm_cache.emplace(k, cg);
    rt->trim_caches();
    rt->gemm_plan_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rt->n_gemm_plans++;
    return cg;
}

static void launch_mul(const CachedGemm* cg, const DTensor& l, const DTensor& r, DTensor& out, int precision);

// sectors without a matching pair stay zero (Tout.setZero(), TensorBase.cpp:256-258): one launch
// over the plan's merged ranges of unwritten sectors
static void zero_unwritten(Runtime* rt, const CachedGemm& cg, DTensor& out)
{
    if(cg.zero_ranges.empty() || out.size() == 0) return;
    dcheck(rt->ctx, xh_fill_ranges(rt->ctx, out.data(), out.size(), cg.zero_ranges.data(), static_cast<int64_t>(cg.zero_ranges.size() / 2), 0.0),
           "zero fill");
}

// While GEMM launches are timed, contraction products carry tag 1 in the launch log and are
// bracketed in a kernel trace by marker kernels of 3 (before) and 4 (after) workgroups, so that
// per-dispatch counters can be split between contractions and the other grouped GEMM calls.
static void tag_contraction(Runtime* rt, bool begin)
{
    if(!rt->time_gemm) return;
    if(begin) {
        dcheck(rt->ctx, xh_trace_marker(rt->ctx, 3), "trace marker");
        xh_set_gemm_tag(rt->ctx, 1);
    } else {
        xh_set_gemm_tag(rt->ctx, 0);
        dcheck(rt->ctx, xh_trace_marker(rt->ctx, 4), "trace marker");
    }
}

DTensor mul(const DTensor& l, const DTensor& r, int precision)
{
    auto cg = get_gemm_plan(l.rt, *l.st, *r.st);
    DTensor out = DTensor::empty(l.rt, cg->plan.out);
    launch_mul(cg.get(), l, r, out, precision);
    if(Tape* tp = recording(l))
        tp->backs.push_back([l, r, out, precision](Tape& tape) {
            // ADTensor.hpp:745-746: left.adj() += res.adj() * right^dagger, right.adj() += left^dagger * res.adj()
            const DTensor* d = tape.find(out);
            if(!d) return;
            const DTensor dc = *d;
            DTensor dl = tape.adjoint_of(l);
            DTensor dr = tape.adjoint_of(r);
            mul_bwd(l, r, dc, &dl, &dr, precision);
        });
    return out;
}

std::vector<DTensor> mul_batch(const std::vector<std::pair<DTensor, DTensor>>& pairs, int precision)
{
    // operator* of independent pairs in one grouped launch: every pair's items, with their offsets
    // taken from the lowest data address of the left operands, right operands and outputs
    const std::size_t n = pairs.size();
    std::vector<DTensor> outs(n);
    if(n == 0) return outs;
    if(n == 1) {
        outs[0] = mul(pairs[0].first, pairs[0].second, precision);
        return outs;
    }
    Runtime* rt = pairs[0].first.rt;
    std::vector<std::shared_ptr<CachedGemm>> cgs(n);
    for(std::size_t i = 0; i < n; ++i) {
        if(pairs[i].first.rt != rt || pairs[i].second.rt != rt) throw Error(XH_EINVAL, "mul_batch: tensors of different runtimes");
        cgs[i] = get_gemm_plan(rt, *pairs[i].first.st, *pairs[i].second.st);
        outs[i] = DTensor::empty(rt, cgs[i]->plan.out);
    }
    if(rt->ctx) {
        const double *A0 = nullptr, *B0 = nullptr;
        double* C0 = nullptr;
        for(std::size_t i = 0; i < n; ++i) {
            if(cgs[i]->plan.items.empty()) continue;
            const double *a = pairs[i].first.data(), *b = pairs[i].second.data();
            double* c = outs[i].data();
            if(!A0 || a < A0) A0 = a;
            if(!B0 || b < B0) B0 = b;
            if(!C0 || c < C0) C0 = c;
        }
        std::vector<xh_gemm_item> items;
        double flops = 0.0;
        for(std::size_t i = 0; i < n; ++i) {
            const CachedGemm& cg = *cgs[i];
            zero_unwritten(rt, cg, outs[i]);
            if(cg.plan.items.empty()) continue;
            const int64_t da = pairs[i].first.data() - A0, db = pairs[i].second.data() - B0, dc = outs[i].data() - C0;
            for(const auto& g : cg.plan.items) {
                items.push_back(xh_gemm_item{g.a_off + da, g.b_off + db, g.c_off + dc, g.m, g.n, g.k});
                flops += 2.0 * g.m * double(g.n) * g.k;
                if(rt->time_gemm) rt->gemm_log.push_back({static_cast<int32_t>(g.m), static_cast<int32_t>(g.n), static_cast<int32_t>(g.k)});
            }
        }
        if(!items.empty()) {
            xh_gemm_plan* plan = nullptr;
            dcheck(rt->ctx, xh_gemm_plan_create(rt->ctx, items.data(), static_cast<int64_t>(items.size()), 0, 0, &plan), "batched gemm plan");
            tag_contraction(rt, true);
            const int rc = xh_gemm(rt->ctx, plan, A0, B0, C0, 1.0, 0.0, precision);
            tag_contraction(rt, false);
            xh_gemm_plan_destroy(rt->ctx, plan);  // its tables are freed after the launch
            dcheck(rt->ctx, rc, "batched grouped gemm");
            rt->gemm_flops += flops;
            rt->n_gemm++;
        }
    }
    for(std::size_t i = 0; i < n; ++i)
        if(Tape* tp = recording(pairs[i].first))
            tp->backs.push_back([l = pairs[i].first, r = pairs[i].second, out = outs[i], precision](Tape& tape) {
                const DTensor* d = tape.find(out);
                if(!d) return;
                const DTensor dc = *d;
                DTensor dl = tape.adjoint_of(l);
                DTensor dr = tape.adjoint_of(r);
                mul_bwd(l, r, dc, &dl, &dr, precision);
            });
    return outs;
}

std::vector<DTensor> contract_batch(const std::vector<ContractJob>& jobs, int precision)
{
    // Tensor::contract of independent jobs: their permutes one by one, the products in one launch
    std::vector<std::pair<DTensor, DTensor>> pairs;
    std::vector<ContractPerms> perms;
    for(const ContractJob& j : jobs) {
        if(static_cast<int>(j.a1.size()) != j.a.st->rank + j.a.st->corank || static_cast<int>(j.a2.size()) != j.b.st->rank + j.b.st->corank)
            throw Error(XH_EINVAL, "contract: index list length does not match the tensor rank");
        ContractPerms cp = get_permutations(j.a1, j.a.st->rank, j.a2, j.b.st->rank, j.rank);
        pairs.emplace_back(permute(j.a, cp.shift1, cp.p1), permute(j.b, cp.shift2, cp.p2));
        perms.push_back(cp);
    }
    std::vector<DTensor> prods = mul_batch(pairs, precision);
    std::vector<DTensor> out;
    for(std::size_t i = 0; i < jobs.size(); ++i) out.push_back(permute(prods[i], perms[i].shiftres, perms[i].pres));
    return out;
}

void mul_into(const DTensor& l, const DTensor& r, DTensor& out, int precision)
{
    auto cg = get_gemm_plan(l.rt, *l.st, *r.st);
    if(out.rt != l.rt || !out.st || (out.st != cg->plan.out && structure_key(*out.st) != structure_key(*cg->plan.out)))
        throw Error(XH_EINVAL, "mul_into: the output tensor does not have the structure of the product");
    launch_mul(cg.get(), l, r, out, precision);
}

static void launch_mul(const CachedGemm* cg, const DTensor& l, const DTensor& r, DTensor& out, int precision)
{
    if(l.rt->ctx && out.size() > 0) {
        Runtime* rt = l.rt;
        zero_unwritten(rt, *cg, out);
        tag_contraction(rt, true);
        const int rc = xh_gemm(rt->ctx, cg->dev, l.data(), r.data(), out.data(), 1.0, 0.0, precision);
        tag_contraction(rt, false);
        dcheck(rt->ctx, rc, "grouped gemm");
        if(rt->time_gemm)
            for(const auto& g : cg->plan.items) rt->gemm_log.push_back({static_cast<int32_t>(g.m), static_cast<int32_t>(g.n), static_cast<int32_t>(g.k)});
        rt->gemm_flops += cg->plan.flops;
        rt->n_gemm++;
    }
}

// The permute of contract()'s left operand fused into the product (CachedGather), built on the
// first use of (left structure, shift, p, right structure).  It qualifies when every GEMM item is a
// skinny FP64 item with the permuted operand on its long side and the permute is single-source
// with coefficients +-1 (abelian symmetries); otherwise dev stays null.
static std::shared_ptr<CachedGather> get_gather(Runtime* rt, const TensorStructure& in, int shift, const std::vector<int>& p, const TensorStructure& right)
{
    std::string k = structure_key(in);
    put_i(k, shift);
    for(int x : p) put_i(k, x);
    k += "|";
    k += structure_key(right);
    auto it = rt->gather_cache.find(k);
    if(it != rt->gather_cache.end()) {
        it->second->last_use = ++rt->tick;
        return it->second;
    }
    auto t0 = std::chrono::steady_clock::now();
    auto tl = t0;  // end of the last lap
    auto lap = [rt, &tl](int slot) {
        const auto now = std::chrono::steady_clock::now();
        rt->host_t[slot] += std::chrono::duration<double>(now - tl).count();
        tl = now;
    };
    double own_s = 0.0;  // this plan's host time, without the GEMM plan's (counted there)
    auto g = std::make_shared<CachedGather>();
    g->last_use = ++rt->tick;
    StructPtr out = permuted_structure(in, shift, p);
    {
        // share the structure with the permute path (same leg sets give the same structure)
        std::string sk = legs_key(out->sym, out->unc_dom, out->unc_cod);
        auto si = rt->struct_cache.find(sk);
        if(si == rt->struct_cache.end())
            rt->struct_cache.emplace(sk, out);
        else if(si->second->sector == out->sector)
            out = si->second;
    }
    g->permuted = out;
    lap(5);
    own_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    g->gemm = get_gemm_plan(rt, *out, right);
    t0 = tl = std::chrono::steady_clock::now();
    int ok = 0;
    if(rt->ctx && g->gemm->dev) dcheck(rt->ctx, xh_gemm_gather_eligible(g->gemm->dev, &ok), "gather eligibility");
    if(ok) {
        std::shared_ptr<const SymbolicPermute> sp = symbolic_permute(rt, in, *out, shift, p);
        lap(6);
        PermutePlan pl = numeric_permute_plan(*sp, in, out, p);
        lap(0);
        for(const PermuteTask& t : pl.tasks) {
            if(t.nsrc != 1) {
                ok = 0;
                break;
            }
            const double c = pl.sources[t.src_begin].coeff;
            if(c != 1.0 && c != -1.0) {
                ok = 0;
                break;
            }
        }
        if(ok) {
            // per GEMM item: the permuted sector that is its A block, its task grid and tree starts
            std::unordered_map<std::int64_t, std::size_t> sector_at;
            for(std::size_t s = 0; s < out->sector.size(); ++s)
                if(out->rows[s] > 0 && out->cols[s] > 0) sector_at.emplace(static_cast<std::int64_t>(out->offsets[s]), s);
            const auto& items = g->gemm->plan.items;
            std::vector<xh_gather_item> gi(items.size(), xh_gather_item{0, 0, 0, 0, 0});
            std::vector<std::int32_t> starts;
            for(std::size_t j = 0; j < items.size(); ++j) {
                if(items[j].m == 0 || items[j].n == 0) continue;
                auto f = sector_at.find(items[j].a_off);
                if(f == sector_at.end()) throw Error(XH_EINVAL, "gather plan: a GEMM item's A block is not a sector of the permuted tensor");
                const std::size_t s = f->second;
                const auto& dt = out->dom().tree(out->sector[s]);
                const auto& ct = out->cod().tree(out->sector[s]);
                gi[j] = xh_gather_item{sp->task_base[s], static_cast<std::int32_t>(dt.size()), static_cast<std::int32_t>(ct.size()), 0,
                                       static_cast<std::int64_t>(starts.size())};
                std::int64_t acc = 0;
                starts.push_back(0);
                for(const auto& t : dt) starts.push_back(static_cast<std::int32_t>(acc += static_cast<std::int64_t>(t.dim)));
                acc = 0;
                starts.push_back(0);
                for(const auto& t : ct) starts.push_back(static_cast<std::int32_t>(acc += static_cast<std::int64_t>(t.dim)));
            }
            xh_gather_desc desc{};
            desc.perm = xh_permute_desc{pl.nlegs, pl.new_rank, static_cast<int64_t>(pl.tasks.size()), static_cast<int64_t>(pl.sources.size()),
                                        reinterpret_cast<const xh_permute_task*>(pl.tasks.data()), reinterpret_cast<const xh_permute_source*>(pl.sources.data())};
            desc.nitems = static_cast<int64_t>(gi.size());
            desc.items = gi.data();
            desc.nstarts = static_cast<int64_t>(starts.size());
            desc.starts = starts.data();
            desc.src_elems = static_cast<int64_t>(in.size);
            lap(1);
            dcheck(rt->ctx, xh_gather_plan_create(rt->ctx, g->gemm->dev, &desc, &g->dev), "gather plan");
            lap(2);
        }
        recycle_plan_tables(pl);
    }
    rt->gather_cache.emplace(k, g);
    rt->trim_caches();
    own_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rt->permute_plan_s += own_s;
    rt->n_permute_plans++;
    return g;
}

// The permute of a gather product folded into its stores (CachedGather::Scatter), built on the
// first use of (shift, p) for that product.  It qualifies when the permute is single-source with
// coefficients +-1 and every sector of the product is written by a GEMM item.
static const CachedGather::Scatter& get_scatter(Runtime* rt, CachedGather& g, int shift, const std::vector<int>& p)
{
    std::string k;
    put_i(k, shift);
    for(int x : p) put_i(k, x);
    auto it = g.scatter.find(k);
    if(it != g.scatter.end()) return it->second;
    const auto t0 = std::chrono::steady_clock::now();
    CachedGather::Scatter sc;
    const TensorStructure& prod = *g.gemm->plan.out;
    StructPtr out = permuted_structure(prod, shift, p);
    {
        std::string sk = legs_key(out->sym, out->unc_dom, out->unc_cod);
        auto si = rt->struct_cache.find(sk);
        if(si == rt->struct_cache.end())
            rt->struct_cache.emplace(sk, out);
        else if(si->second->sector == out->sector)
            out = si->second;
    }
    sc.out = out;
    // every product sector must be some item's output (no zero-filled sectors, no empty items)
    const auto& items = g.gemm->plan.items;
    std::unordered_map<std::int64_t, std::size_t> item_at;  // product sector offset -> GEMM item
    bool ok = g.dev != nullptr && g.gemm->zero_rects.empty();
    for(std::size_t j = 0; ok && j < items.size(); ++j) {
        if(items[j].m == 0 || items[j].n == 0) ok = false;
        item_at.emplace(items[j].c_off, j);
    }
    std::vector<std::size_t> sector_item(prod.sector.size());
    for(std::size_t s = 0; ok && s < prod.sector.size(); ++s) {
        auto f = item_at.find(static_cast<std::int64_t>(prod.offsets[s]));
        if(f == item_at.end() || prod.rows[s] == 0 || prod.cols[s] == 0)
            ok = false;
        else
            sector_item[s] = f->second;
    }
    if(ok) {
        std::shared_ptr<const SymbolicPermute> sp = symbolic_permute(rt, prod, *out, shift, p);
        PermutePlan pl = numeric_permute_plan(*sp, prod, out, p);
        for(const PermuteTask& t : pl.tasks) {
            const double c = t.nsrc == 1 ? pl.sources[t.src_begin].coeff : 0.0;
            if(t.nsrc != 1 || (c != 1.0 && c != -1.0)) {
                ok = false;
                break;
            }
        }
        // per item the task grid of its product block and its column starts
        std::vector<std::int64_t> grid_off(items.size(), 0), cstarts_off(items.size(), 0);
        std::vector<std::int32_t> ncod(items.size(), 0), grid, cstarts;
        if(ok) {
            for(std::size_t s = 0; s < prod.sector.size(); ++s) {
                const std::size_t j = sector_item[s];
                const auto& dt = prod.dom().tree(prod.sector[s]);
                const auto& ct = prod.cod().tree(prod.sector[s]);
                grid_off[j] = static_cast<std::int64_t>(grid.size());
                ncod[j] = static_cast<std::int32_t>(ct.size());
                grid.resize(grid.size() + dt.size() * ct.size(), -1);
                cstarts_off[j] = static_cast<std::int64_t>(cstarts.size());
                std::int64_t acc = 0;
                cstarts.push_back(0);
                for(const auto& t : ct) cstarts.push_back(static_cast<std::int32_t>(acc += static_cast<std::int64_t>(t.dim)));
            }
            for(std::size_t t = 0; t < pl.tasks.size(); ++t) {
                const auto& src = sp->sources[static_cast<std::size_t>(sp->src_begin[t])];
                const std::size_t j = sector_item[static_cast<std::size_t>(src.sector)];
                grid[static_cast<std::size_t>(grid_off[j] + static_cast<std::int64_t>(src.t1) * ncod[j] + src.t2)] = static_cast<std::int32_t>(t);
            }
            for(std::int32_t t : grid)
                if(t < 0) ok = false;  // a product subblock that no task reads: leave the permute to permute()
        }
        if(ok) {
            xh_scatter_desc desc{};
            desc.perm = xh_permute_desc{pl.nlegs, pl.new_rank, static_cast<int64_t>(pl.tasks.size()), static_cast<int64_t>(pl.sources.size()),
                                        reinterpret_cast<const xh_permute_task*>(pl.tasks.data()), reinterpret_cast<const xh_permute_source*>(pl.sources.data())};
            std::vector<std::int32_t> p32(p.begin(), p.end());
            desc.p = p32.data();
            desc.nitems = static_cast<int64_t>(items.size());
            desc.grid_off = grid_off.data();
            desc.ncod = ncod.data();
            desc.cstarts_off = cstarts_off.data();
            desc.ngrid = static_cast<int64_t>(grid.size());
            desc.grid = grid.data();
            desc.ncstarts = static_cast<int64_t>(cstarts.size());
            desc.cstarts = cstarts.data();
            desc.dst_elems = static_cast<int64_t>(out->size);
            dcheck(rt->ctx, xh_scatter_plan_create(rt->ctx, g.dev, &desc, &sc.dev), "scatter plan");
        }
        recycle_plan_tables(pl);
    }
    rt->permute_plan_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rt->n_permute_plans++;
    return g.scatter.emplace(k, sc).first->second;
}

// the product of a gather plan with its permute folded in: out = permute(permute(a) * pb)
static void launch_gather_scatter(const CachedGather* g, const CachedGather::Scatter& sc, const DTensor& a, const DTensor& pb, DTensor& out)
{
    Runtime* rt = a.rt;
    if(out.size() == 0) return;
    const CachedGemm* cg = g->gemm.get();
    tag_contraction(rt, true);
    const int rc = xh_gemm_gather_scatter(rt->ctx, cg->dev, g->dev, sc.dev, a.data(), static_cast<int64_t>(a.size()), pb.data(), out.data(),
                                          static_cast<int64_t>(out.size()), 1.0);
    tag_contraction(rt, false);
    dcheck(rt->ctx, rc, "gather gemm with the product's permute");
    if(rt->time_gemm)
        for(const auto& it : cg->plan.items) rt->gemm_log.push_back({static_cast<int32_t>(it.m), static_cast<int32_t>(it.n), static_cast<int32_t>(it.k)});
    rt->gemm_flops += cg->plan.flops;
    rt->n_gemm++;
    rt->n_fused++;
}

// the product of a gather plan: out = permute(a) * pb with permute(a) read inside the GEMM
static void launch_gather(const CachedGather* g, const DTensor& a, const DTensor& pb, DTensor& out)
{
    Runtime* rt = a.rt;
    if(out.size() == 0) return;
    const CachedGemm* cg = g->gemm.get();
    zero_unwritten(rt, *cg, out);  // sectors without a matching pair stay zero, as in operator*
    tag_contraction(rt, true);
    const int rc = xh_gemm_gather(rt->ctx, cg->dev, g->dev, a.data(), static_cast<int64_t>(a.size()), pb.data(), out.data(), 1.0, 0.0);
    tag_contraction(rt, false);
    dcheck(rt->ctx, rc, "gather gemm");
    if(rt->time_gemm)
        for(const auto& it : cg->plan.items) rt->gemm_log.push_back({static_cast<int32_t>(it.m), static_cast<int32_t>(it.n), static_cast<int32_t>(it.k)});
    rt->gemm_flops += cg->plan.flops;
    rt->n_gemm++;
    rt->n_fused++;
}

DTensor contract(const DTensor& a, const std::vector<int>& a1, const DTensor& b, const std::vector<int>& a2, int rankres, int precision)
{
    // Tensor::contract (Tensor.hpp:328-351)
    if(static_cast<int>(a1.size()) != a.st->rank + a.st->corank || static_cast<int>(a2.size()) != b.st->rank + b.st->corank)
        throw Error(XH_EINVAL, "contract: index list length does not match the tensor rank");
    ContractPerms cp = get_permutations(a1, a.st->rank, a2, b.st->rank, rankres);
    Runtime* rt = a.rt;
    bool ident = cp.shift1 == 0;
    for(std::size_t i = 0; ident && i < cp.p1.size(); ++i) ident = cp.p1[i] == static_cast<int>(i);
    if(rt->fuse_permute > 0 && rt->ctx && precision == XH_FP64 && !ident && a.size() > 0 && b.rt == rt && !recording(a) && !recording(b)) {
        // the left permute read inside the product when the product qualifies (xh_gemm_gather)
        DTensor pb = permute(b, cp.shift2, cp.p2);
        auto g = get_gather(rt, *a.st, cp.shift1, cp.p1, *pb.st);
        if(g->dev) {
            bool rid = cp.shiftres == 0;
            for(std::size_t i = 0; rid && i < cp.pres.size(); ++i) rid = cp.pres[i] == static_cast<int>(i);
            if(!rid && rt->fuse_permute >= 2) {
                const CachedGather::Scatter& sc = get_scatter(rt, *g, cp.shiftres, cp.pres);
                if(sc.dev) {
                    DTensor out = DTensor::empty(rt, sc.out);
                    launch_gather_scatter(g.get(), sc, a, pb, out);
                    return out;
                }
            }
            DTensor prod = DTensor::empty(rt, g->gemm->plan.out);
            launch_gather(g.get(), a, pb, prod);
            return permute(prod, cp.shiftres, cp.pres);
        }
        DTensor pa = permute(a, cp.shift1, cp.p1);
        return permute(mul(pa, pb, precision), cp.shiftres, cp.pres);
    }
    DTensor pa = permute(a, cp.shift1, cp.p1);
    DTensor pb = permute(b, cp.shift2, cp.p2);
    DTensor prod = mul(pa, pb, precision);
    return permute(prod, cp.shiftres, cp.pres);
}

// ------------------------------------------------------------------ elementwise

DTensor adjoint(const DTensor& t)
{
    // AdjointOp + Tensor(TensorBase) (AdjointOp.hpp:25-69, Tensor.hpp:446-452): bases swapped
    // without conjugation, sectors in the original order, every block transposed.
    auto st = TensorStructure::create_pushed(t.st->sym, t.st->unc_cod, t.st->unc_dom, t.st->sector);
    DTensor out = DTensor::empty(t.rt, st);
    if(t.rt->ctx && out.size() > 0) {
        std::vector<xh_copy_rect> rects;
        for(std::size_t s = 0; s < st->sector.size(); ++s) {
            // out block s is cols_in x rows_in
            rects.push_back(xh_copy_rect{static_cast<int64_t>(t.st->offsets[s]), static_cast<int64_t>(t.st->rows[s]), static_cast<int64_t>(st->offsets[s]),
                                         static_cast<int64_t>(st->rows[s]), static_cast<int32_t>(st->rows[s]), static_cast<int32_t>(st->cols[s]), 1.0, 1});
        }
        dcheck(t.rt->ctx, xh_copy_rects(t.rt->ctx, t.data(), out.data(), rects.data(), static_cast<int64_t>(rects.size())), "adjoint");
    }
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out](Tape& tape) {
            // ADTensor.hpp:138-150
            if(const DTensor* d = tape.find(out)) tape.accumulate(t, adjoint(*d));
        });
    return out;
}

static DTensor copy_of(const DTensor& t)
{
    DTensor out = DTensor::empty(t.rt, t.st);
    if(t.rt->ctx && out.size() > 0) dcheck(t.rt->ctx, xh_axpby(t.rt->ctx, out.data(), t.data(), t.size(), 1.0, 0.0), "copy");
    return out;
}

DTensor twist(const DTensor& t, int leg)
{
    // Tensor::twist (Tensor.cpp:357-377): identity for bosonic symmetries
    if(!t.st->sym.any_fermionic()) return t;
    const TensorStructure& s = *t.st;
    if(leg < 0 || leg >= s.rank + s.corank) throw Error(XH_EINVAL, "twist: leg out of range");
    DTensor out = copy_of(t);
    std::vector<xh_rect> rects;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        std::size_t ro = 0;
        for(const auto& t1 : s.dom().tree(s.sector[i])) {
            std::size_t co = 0;
            for(const auto& t2 : s.cod().tree(s.sector[i])) {
                double c = leg < s.rank ? s.sym.coeff_twist(t1.unc[leg]) : s.sym.coeff_twist(t2.unc[leg - s.rank]);
                if(c != 1.0)
                    rects.push_back(xh_rect{static_cast<int64_t>(s.offsets[i] + ro + co * s.rows[i]), static_cast<int64_t>(s.rows[i]),
                                            static_cast<int32_t>(t1.dim), static_cast<int32_t>(t2.dim), c});
                co += t2.dim;
            }
            ro += t1.dim;
        }
    }
    if(t.rt->ctx && !rects.empty()) dcheck(t.rt->ctx, xh_scale_rects(t.rt->ctx, out.data(), rects.data(), static_cast<int64_t>(rects.size())), "twist");
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out, leg](Tape& tape) {
            // ADTensor.hpp:592-602: the twist is a diagonal sign map, its own transpose
            if(const DTensor* d = tape.find(out)) tape.accumulate(t, twist(*d, leg));
        });
    return out;
}

DTensor scaled(const DTensor& t, double alpha)
{
    DTensor out = DTensor::empty(t.rt, t.st);
    if(t.rt->ctx && out.size() > 0) dcheck(t.rt->ctx, xh_axpby(t.rt->ctx, out.data(), t.data(), t.size(), alpha, 0.0), "scale");
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out, alpha](Tape& tape) {
            if(const DTensor* d = tape.find(out)) tape.accumulate(t, *d, alpha);
        });
    return out;
}

DTensor axpby(const DTensor& x, const DTensor& y, double a, double b)
{
    if(structure_key(*x.st) != structure_key(*y.st)) throw Error(XH_EINVAL, "binary op: tensors have different structure");
    DTensor out = copy_of(y);
    if(x.rt->ctx && out.size() > 0) dcheck(x.rt->ctx, xh_axpby(x.rt->ctx, out.data(), x.data(), x.size(), a, b), "axpby");
    if(Tape* tp = recording(x))
        tp->backs.push_back([x, y, out, a, b](Tape& tape) {
            const DTensor* d = tape.find(out);
            if(!d) return;
            const DTensor dd = *d;
            tape.accumulate(x, dd, a);
            tape.accumulate(y, dd, b);
        });
    return out;
}

DTensor diag_map(const DTensor& t, int mode)
{
    DTensor out = copy_of(t);
    std::vector<int64_t> off;
    std::vector<int32_t> m;
    for(std::size_t i = 0; i < t.st->sector.size(); ++i) {
        if(t.st->rows[i] != t.st->cols[i]) throw Error(XH_EINVAL, "diagonal map needs square blocks");
        off.push_back(static_cast<int64_t>(t.st->offsets[i]));
        m.push_back(static_cast<int32_t>(t.st->rows[i]));
    }
    // the reference's result is the diagonal matrix of the mapped diagonal (diagUnaryFunc)
    if(t.rt->ctx && !off.empty())
        dcheck(t.rt->ctx, xh_block_diag_map(t.rt->ctx, off.data(), m.data(), static_cast<int64_t>(off.size()), out.data(), mode | 4), "diag map");
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out, mode](Tape& tape) {
            // ADTensor.hpp:560-590 (diagonal entries only)
            const DTensor* d = tape.find(out);
            if(!d) return;
            const DTensor dd = *d;
            DTensor dt = tape.adjoint_of(t);
            diag_bwd(out, dd, dt, mode);
        });
    return out;
}

DTensor identity(Runtime* rt, const std::vector<Qbasis>& dom, const std::vector<Qbasis>& cod)
{
    // Tensor::Identity / setIdentity (Tensor.hpp:229-235, Tensor.cpp:58-68): every block gets ones
    // on its main diagonal; the pattern is written on the host and uploaded
    if(dom.empty() && cod.empty()) throw Error(XH_EINVAL, "identity needs at least one leg");
    StructPtr st = cached_structure(rt, dom.empty() ? cod[0].sym : dom[0].sym, dom, cod);
    DTensor t = DTensor::empty(rt, st);
    std::vector<double> h(st->size, 0.0);
    for(std::size_t i = 0; i < st->sector.size(); ++i)
        for(std::size_t k = 0; k < std::min(st->rows[i], st->cols[i]); ++k) h[st->offsets[i] + k + k * st->rows[i]] = 1.0;
    if(st->size) t.upload(h.data(), h.size());
    return t;
}

DTensor mexp(const DTensor& t, double factor)
{
    t.rt->require_device();
    DTensor out = DTensor::empty(t.rt, t.st);
    std::vector<xh_mexp_item> items;
    for(std::size_t i = 0; i < t.st->sector.size(); ++i) {
        if(t.st->rows[i] != t.st->cols[i]) throw Error(XH_EINVAL, "mexp needs square blocks");
        if(t.st->rows[i] > XH_MEXP_MAX) throw Error(XH_EINVAL, "mexp: blocks larger than 64 x 64 are not supported");
        items.push_back(xh_mexp_item{static_cast<int64_t>(t.st->offsets[i]), static_cast<int64_t>(t.st->offsets[i]), static_cast<int32_t>(t.st->rows[i]), 0});
    }
    if(!items.empty()) dcheck(t.rt->ctx, xh_block_mexp(t.rt->ctx, items.data(), static_cast<int64_t>(items.size()), t.data(), out.data(), factor), "mexp");
    return out;
}

double trace(const DTensor& t)
{
    t.rt->require_device();
    std::vector<int64_t> off;
    std::vector<int32_t> m;
    for(std::size_t i = 0; i < t.st->sector.size(); ++i) {
        if(t.st->rows[i] != t.st->cols[i]) throw Error(XH_EINVAL, "trace needs square blocks");
        off.push_back(static_cast<int64_t>(t.st->offsets[i]));
        m.push_back(static_cast<int32_t>(t.st->rows[i]));
    }
    std::vector<double> tr(off.size());
    if(!off.empty()) dcheck(t.rt->ctx, xh_block_trace(t.rt->ctx, off.data(), m.data(), static_cast<int64_t>(off.size()), t.data(), tr.data()), "trace");
    double out = 0.0;
    for(std::size_t i = 0; i < tr.size(); ++i) out += tr[i] * static_cast<double>(t.st->sym.degeneracy(t.st->sector[i]));
    return out;
}

double squared_norm(const DTensor& t)
{
    // (T * T^dagger).trace(): sum over blocks of degeneracy * |block|_F^2
    t.rt->require_device();
    double out = 0.0;
    for(std::size_t i = 0; i < t.st->sector.size(); ++i) {
        double d = 0;
        const std::size_t n = t.st->rows[i] * t.st->cols[i];
        if(n) dcheck(t.rt->ctx, xh_dot(t.rt->ctx, t.data() + t.st->offsets[i], t.data() + t.st->offsets[i], n, &d), "squaredNorm");
        out += d * static_cast<double>(t.st->sym.degeneracy(t.st->sector[i]));
    }
    return out;
}

double max_norm(const DTensor& t)
{
    t.rt->require_device();
    double out = 0.0;
    if(t.size()) dcheck(t.rt->ctx, xh_max_abs(t.rt->ctx, t.data(), t.size(), &out), "maxNorm");
    return out;
}

// ------------------------------------------------------------------ tSVD

TruncRule truncation_rule(const Sym& s, std::vector<std::pair<Q, double>> spectrum, std::size_t max_keep, double eps_svd, bool preserve_multiplets,
                          double trunc_weight, double entropy)
{
    // The selection of cutoff_matrices (Tensor.cpp:535-598).  The candidates are ordered by
    // std::sort with a descending-value comparator, as there, so that equal values end up in
    // libstdc++'s introsort order and the same ones are kept.
    using Entry = std::pair<Q, double>;
    if(spectrum.empty()) throw Error(XH_EINVAL, "tSVD of a tensor without sectors");
    // std::sort needs a strict weak order: NaN (from non-finite input data) would break it
    for(const Entry& e : spectrum)
        if(!std::isfinite(e.second)) throw Error(XH_EINVAL, "tSVD: non-finite singular value (non-finite tensor data)");
    const Entry fallback = spectrum.front();  // kept alone when every value is cut
    const std::size_t n_all = spectrum.size();
    std::sort(spectrum.begin(), spectrum.end(), [](const Entry& a, const Entry& b) { return a.second > b.second; });
    // discarded weight: every value past the first max_keep, times its multiplet degeneracy
    for(std::size_t k = max_keep; k < n_all; ++k) trunc_weight += s.degeneracy(spectrum[k].first) * std::pow(std::abs(spectrum[k].second), 2.);
    spectrum.resize(std::min(max_keep, n_all));
    if(eps_svd > 0.) {
        auto below = [eps_svd](const Entry& e) { return e.second < eps_svd; };
        spectrum.erase(std::remove_if(spectrum.begin(), spectrum.end(), below), spectrum.end());
    }
    if(preserve_multiplets) {
        // cut after the smallest-value position whose relative gap to the value above exceeds 10 %
        for(std::size_t k = spectrum.size(); k-- > 1;) {
            const double above = spectrum[k - 1].second, here = spectrum[k].second;
            if(2 * (above - here) / (above + here) > 0.1) {
                spectrum.resize(k);
                break;
            }
        }
    }
    if(spectrum.empty()) spectrum.push_back(fallback);
    TruncRule out;
    for(const auto& [q, sigma] : spectrum) {
        out.kept[q].push_back(sigma);
        entropy += -s.degeneracy(q) * sigma * sigma * std::log(sigma * sigma);
    }
    out.trunc_weight = trunc_weight;
    out.entropy = entropy;
    return out;
}

FullSVD full_svd(const DTensor& t)
{
    t.rt->require_device();
    const TensorStructure& s = *t.st;
    Qbasis middle(s.sym, 1);
    for(std::size_t i = 0; i < s.sector.size(); ++i) middle.push_back(s.sector[i], std::min(s.rows[i], s.cols[i]));
    FullSVD f;
    f.sectors = s.sector;
    // U(uncoupledDomain, {middle}), Sigma({middle},{middle}), Vdag({middle}, uncoupledCodomain), filled by push_back
    f.U = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {middle}, s.sector));
    f.S = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {middle}, {middle}, s.sector));
    f.Vdag = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {middle}, s.unc_cod, s.sector));
    std::vector<xh_factor_item> items;
    std::size_t nsv = 0;
    std::vector<std::size_t> sv_off;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        sv_off.push_back(nsv);
        nsv += std::min(s.rows[i], s.cols[i]);
    }
    double* d_sv = nullptr;
    if(nsv) dcheck(t.rt->ctx, xh_alloc(t.rt->ctx, nsv * sizeof(double), reinterpret_cast<void**>(&d_sv)), "alloc");
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        items.push_back(xh_factor_item{static_cast<int64_t>(s.offsets[i]), static_cast<int64_t>(f.U.st->offsets[i]), static_cast<int64_t>(sv_off[i]),
                                       static_cast<int64_t>(f.Vdag.st->offsets[i]), static_cast<int32_t>(s.rows[i]), static_cast<int32_t>(s.cols[i])});
    }
    if(!items.empty()) dcheck(t.rt->ctx, xh_block_svd(t.rt->ctx, items.data(), static_cast<int64_t>(items.size()), t.data(), f.U.data(), d_sv, f.Vdag.data()), "block svd");
    std::vector<double> sv(nsv);
    if(nsv) dcheck(t.rt->ctx, xh_download(t.rt->ctx, sv.data(), d_sv, nsv * sizeof(double)), "download singular values");
    if(d_sv) xh_free(t.rt->ctx, d_sv);
    // Sigma blocks: diagonal matrices
    std::vector<double> hs(f.S.size(), 0.0);
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        const std::size_t r = std::min(s.rows[i], s.cols[i]);
        for(std::size_t k = 0; k < r; ++k) {
            hs[f.S.st->offsets[i] + k + k * r] = sv[sv_off[i] + k];
            f.all_sv.emplace_back(s.sector[i], sv[sv_off[i] + k]);
        }
    }
    if(f.S.size()) f.S.upload(hs.data(), hs.size());
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, U = f.U, S = f.S, V = f.Vdag](Tape& tape) {
            // ADTensor.hpp:262-320
            if(!tape.find(U) && !tape.find(S) && !tape.find(V)) return;
            const DTensor dU = tape.adjoint_of(U), dS = tape.adjoint_of(S), dV = tape.adjoint_of(V);
            DTensor dt = tape.adjoint_of(t);
            svd_bwd(U, S, V, dU, dS, dV, dt);
        });
    return f;
}

SVDResult cutoff(const DTensor& t, const FullSVD& f, std::size_t maxKeep, double eps_svd, bool preserve_multiplets, double trunc_weight_in,
                 double entropy_in)
{
    const TensorStructure& s = *t.st;
    TruncRule rule = truncation_rule(s.sym, f.all_sv, maxKeep, eps_svd, preserve_multiplets, trunc_weight_in, entropy_in);
    Qbasis truncBasis(s.sym, 1);
    std::vector<Q> order;
    for(const auto& [q, v] : rule.kept) {
        truncBasis.push_back(q, v.size());
        order.push_back(q);
    }
    SVDResult res;
    res.trunc_weight = rule.trunc_weight;
    res.entropy = rule.entropy;
    res.spec = rule.kept;
    res.n_states = f.all_sv.size();
    res.U = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {truncBasis}, order));
    res.S = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {truncBasis}, {truncBasis}, order));
    res.Vdag = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {truncBasis}, s.unc_cod, order));
    std::vector<xh_copy_rect> ru, rv;
    std::vector<double> hs(res.S.size(), 0.0);
    for(std::size_t o = 0; o < order.size(); ++o) {
        const Q& q = order[o];
        const int i = s.find_sector(q);
        const std::size_t nret = rule.kept[q].size();
        const std::size_t r = std::min(s.rows[i], s.cols[i]);
        ru.push_back(xh_copy_rect{static_cast<int64_t>(f.U.st->offsets[i]), static_cast<int64_t>(s.rows[i]), static_cast<int64_t>(res.U.st->offsets[o]),
                                  static_cast<int64_t>(s.rows[i]), static_cast<int32_t>(s.rows[i]), static_cast<int32_t>(nret), 1.0, 0});
        rv.push_back(xh_copy_rect{static_cast<int64_t>(f.Vdag.st->offsets[i]), static_cast<int64_t>(r), static_cast<int64_t>(res.Vdag.st->offsets[o]),
                                  static_cast<int64_t>(nret), static_cast<int32_t>(nret), static_cast<int32_t>(s.cols[i]), 1.0, 0});
        for(std::size_t k = 0; k < nret; ++k) hs[res.S.st->offsets[o] + k + k * nret] = rule.kept[q][k];
    }
    if(!ru.empty()) {
        dcheck(t.rt->ctx, xh_copy_rects(t.rt->ctx, f.U.data(), res.U.data(), ru.data(), static_cast<int64_t>(ru.size())), "truncate U");
        dcheck(t.rt->ctx, xh_copy_rects(t.rt->ctx, f.Vdag.data(), res.Vdag.data(), rv.data(), static_cast<int64_t>(rv.size())), "truncate Vdag");
    }
    if(res.S.size()) res.S.upload(hs.data(), hs.size());
    if(Tape* tp = recording(t))
        tp->backs.push_back([U = f.U, S = f.S, V = f.Vdag, Ut = res.U, St = res.S, Vt = res.Vdag](Tape& tape) {
            // ADTensor.hpp:357-370
            if(!tape.find(Ut) && !tape.find(St) && !tape.find(Vt)) return;
            const DTensor dUt = tape.adjoint_of(Ut), dSt = tape.adjoint_of(St), dVt = tape.adjoint_of(Vt);
            DTensor dU = tape.adjoint_of(U), dS = tape.adjoint_of(S), dV = tape.adjoint_of(V);
            cutoff_bwd(dUt, dSt, dVt, dU, dS, dV);
        });
    return res;
}

SVDResult tsvd(const DTensor& t, std::size_t maxKeep, double eps_svd, bool preserve_multiplets, double trunc_weight_in, double entropy_in)
{
    FullSVD f = full_svd(t);
    return cutoff(t, f, maxKeep, eps_svd, preserve_multiplets, trunc_weight_in, entropy_in);
}

std::pair<DTensor, DTensor> tqr(const DTensor& t, bool return_lq)
{
    t.rt->require_device();
    const TensorStructure& s = *t.st;
    if(return_lq)
        for(std::size_t i = 0; i < s.sector.size(); ++i)
            if(s.rows[i] != s.cols[i]) throw Error(XH_EINVAL, "tQR(RETURN_LQ): the adjoint factors fit the tensors of Q and R only for square sector blocks");
    Qbasis middle = s.dom().dim() > s.cod().dim() ? s.cod().forget_history() : s.dom().forget_history();
    DTensor Qt = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {middle}, s.sector));
    DTensor Rt = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {middle}, s.unc_cod, s.sector));
    std::vector<xh_factor_item> items;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        const std::size_t mid = middle.inner_dim(s.sector[i]);
        items.push_back(xh_factor_item{static_cast<int64_t>(s.offsets[i]), static_cast<int64_t>(Qt.st->offsets[i]), static_cast<int64_t>(mid),
                                       static_cast<int64_t>(Rt.st->offsets[i]), static_cast<int32_t>(s.rows[i]), static_cast<int32_t>(s.cols[i])});
    }
    if(!items.empty()) dcheck(t.rt->ctx, xh_block_qr(t.rt->ctx, items.data(), static_cast<int64_t>(items.size()), t.data(), Qt.data(), Rt.data()), "block qr");
    if(return_lq && !items.empty()) {
        // Qmat = Q^T, Rmat = R^T per sector (Tensor.cpp:767-769); square blocks, same offsets
        DTensor Ql = DTensor::empty(t.rt, Qt.st), Rl = DTensor::empty(t.rt, Rt.st);
        std::vector<xh_copy_rect> rq, rr;
        for(std::size_t i = 0; i < s.sector.size(); ++i) {
            const int32_t n = static_cast<int32_t>(s.rows[i]);
            if(n == 0) continue;
            rq.push_back(xh_copy_rect{static_cast<int64_t>(Qt.st->offsets[i]), n, static_cast<int64_t>(Qt.st->offsets[i]), n, n, n, 1.0, 1});
            rr.push_back(xh_copy_rect{static_cast<int64_t>(Rt.st->offsets[i]), n, static_cast<int64_t>(Rt.st->offsets[i]), n, n, n, 1.0, 1});
        }
        dcheck(t.rt->ctx, xh_copy_rects(t.rt->ctx, Qt.data(), Ql.data(), rq.data(), static_cast<int64_t>(rq.size())), "transpose Q");
        dcheck(t.rt->ctx, xh_copy_rects(t.rt->ctx, Rt.data(), Rl.data(), rr.data(), static_cast<int64_t>(rr.size())), "transpose R");
        if(recording(t)) throw Error(XH_EINVAL, "tQR(RETURN_LQ) has no recorded reverse pass; record tQR of the adjoint instead");
        return {Ql, Rl};
    }
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, Qt, Rt](Tape& tape) {
            // the QR adjoint (the reference's AD tensor records no tQR)
            if(!tape.find(Qt) && !tape.find(Rt)) return;
            const DTensor dQ = tape.adjoint_of(Qt), dR = tape.adjoint_of(Rt);
            DTensor dt = tape.adjoint_of(t);
            qr_bwd(t, Qt, Rt, dQ, dR, dt);
        });
    return {Qt, Rt};
}

std::pair<DTensor, DTensor> teigh(const DTensor& t)
{
    t.rt->require_device();
    const TensorStructure& s = *t.st;
    if(s.rank != s.corank) throw Error(XH_EINVAL, "teigh needs Rank == CoRank");
    Qbasis dh = s.dom().forget_history(), ch = s.cod().forget_history();
    DTensor D = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {dh}, {ch}, s.sector));
    DTensor V = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {ch}, s.sector));
    std::vector<xh_factor_item> items;
    std::size_t nev = 0;
    std::vector<std::size_t> ev_off;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        if(s.rows[i] != s.cols[i]) throw Error(XH_EINVAL, "teigh: non-square block");
        ev_off.push_back(nev);
        nev += s.rows[i];
        items.push_back(xh_factor_item{static_cast<int64_t>(s.offsets[i]), static_cast<int64_t>(V.st->offsets[i]), static_cast<int64_t>(ev_off.back()), 0,
                                       static_cast<int32_t>(s.rows[i]), static_cast<int32_t>(s.cols[i])});
    }
    double* d_ev = nullptr;
    if(nev) dcheck(t.rt->ctx, xh_alloc(t.rt->ctx, nev * sizeof(double), reinterpret_cast<void**>(&d_ev)), "alloc");
    if(!items.empty()) dcheck(t.rt->ctx, xh_block_eigh(t.rt->ctx, items.data(), static_cast<int64_t>(items.size()), t.data(), d_ev, V.data()), "block eigh");
    std::vector<double> ev(nev);
    if(nev) dcheck(t.rt->ctx, xh_download(t.rt->ctx, ev.data(), d_ev, nev * sizeof(double)), "download eigenvalues");
    if(d_ev) xh_free(t.rt->ctx, d_ev);
    std::vector<double> hd(D.size(), 0.0);
    for(std::size_t i = 0; i < s.sector.size(); ++i)
        for(std::size_t k = 0; k < s.rows[i]; ++k) hd[D.st->offsets[i] + k + k * s.rows[i]] = ev[ev_off[i] + k];
    if(D.size()) D.upload(hd.data(), hd.size());
    return {D, V};
}

// ------------------------------------------------------------------ reverse-mode recording

const DTensor* Tape::find(const DTensor& x) const
{
    if(!x.buf) return nullptr;
    auto it = adj.find(x.buf.get());
    return it == adj.end() ? nullptr : &it->second.adj;
}

DTensor Tape::adjoint_of(const DTensor& x)
{
    if(!x.buf) throw Error(XH_EINVAL, "adjoint of a tensor without storage");
    auto it = adj.find(x.buf.get());
    if(it != adj.end()) return it->second.adj;
    DTensor z = DTensor::zeros(x.rt, x.st);
    adj.emplace(x.buf.get(), Entry{x, z});
    return z;
}

void Tape::accumulate(const DTensor& x, const DTensor& dx, double alpha)
{
    if(!x.buf) return;
    const TensorStructure &a = *x.st, &b = *dx.st;
    if(a.sector != b.sector || a.rows != b.rows || a.cols != b.cols || a.size != b.size)
        throw Error(XH_EINVAL, "reverse pass: adjoint does not have the block structure of its value");
    auto it = adj.find(x.buf.get());
    if(it == adj.end()) {
        DTensor c = DTensor::empty(x.rt, x.st);
        if(x.rt->ctx && c.size() > 0) dcheck(x.rt->ctx, xh_axpby(x.rt->ctx, c.data(), dx.data(), c.size(), alpha, 0.0), "adjoint");
        adj.emplace(x.buf.get(), Entry{x, c});
        return;
    }
    DTensor& d = it->second.adj;
    if(x.rt->ctx && d.size() > 0) dcheck(x.rt->ctx, xh_axpby(x.rt->ctx, d.data(), dx.data(), d.size(), alpha, 1.0), "adjoint accumulation");
}

void Tape::backward()
{
    active = false;
    for(std::size_t i = backs.size(); i-- > 0;) backs[i](*this);
    backs.clear();
}

void Tape::clear()
{
    active = false;
    backs.clear();
    adj.clear();
}

DTensor div_trace(const DTensor& x, const DTensor& s)
{
    const double tr = trace(s);
    if(!(tr != 0.0) || !std::isfinite(tr)) throw Error(XH_EINVAL, "division by a trace of " + std::to_string(tr));
    DTensor out = DTensor::empty(x.rt, x.st);
    if(x.rt->ctx && out.size() > 0) dcheck(x.rt->ctx, xh_axpby(x.rt->ctx, out.data(), x.data(), x.size(), 1.0 / tr, 0.0), "divide by trace");
    if(Tape* tp = recording(x))
        tp->backs.push_back([x, s, out, tr](Tape& tape) {
            // y = x / tr(s): dx += dy / tr, ds += -<dy, x> / tr^2 * WeightedIdentity
            const DTensor* d = tape.find(out);
            if(!d) return;
            const DTensor dd = *d;
            double dot = 0.0;
            if(x.size()) dcheck(x.rt->ctx, xh_dot(x.rt->ctx, dd.data(), x.data(), x.size(), &dot), "dot");
            tape.accumulate(x, dd, 1.0 / tr);
            DTensor ds = tape.adjoint_of(s);
            trace_bwd(ds, -dot / (tr * tr));
        });
    return out;
}

}  // namespace xa
