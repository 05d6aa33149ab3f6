// C ABI of the tensor layer (include/xped_tensor.h) over the C++ structure engine and device layer.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "../../include/xped_tensor.h"
#include "core/ctm.hpp"
#include "core/simple_update.hpp"
#include "core/storage.hpp"
#include "core/ad.hpp"
#include "core/tensor.hpp"

struct xt_rt {
    xa::Runtime rt;
    std::string err;
    long live = 0;  // tensor and CTM handles made on this runtime and not yet destroyed
    std::unique_ptr<xa::Tape> tape;  // set by xt_tape_begin
};
struct xt_basis {
    xa::Qbasis b;
};
struct xt_tensor {
    xa::DTensor t;
    xt_rt* owner = nullptr;
};
struct xt_spec {
    std::map<xa::Q, std::vector<double>> m;  // ordered by q, as std::map<qarray, VectorType>
};

namespace {

thread_local std::string g_err;

template <typename F>
int run(xt_rt* rt, F&& f)
{
    try {
        f();
        return XH_OK;
    } catch(const xa::Error& e) {
        g_err = e.what();
        if(rt) rt->err = g_err;
        return e.code;
    } catch(const std::bad_alloc&) {
        g_err = "host allocation failed";
        if(rt) rt->err = g_err;
        return XH_ENOMEM;
    } catch(const std::exception& e) {
        g_err = e.what();
        if(rt) rt->err = g_err;
        return XH_EINVAL;
    }
}

xa::Q q_from(const int* q)
{
    xa::Q out;
    if(q) {
        out.v[0] = q[0];
        out.v[1] = q[1];
    }
    return out;
}
void q_to(const xa::Q& q, int* out)
{
    out[0] = q.v[0];
    out[1] = q.v[1];
}

xt_tensor* wrap(xt_rt* owner, xa::DTensor t)
{
    auto* h = new xt_tensor();
    h->t = std::move(t);
    h->owner = owner;
    if(owner) ++owner->live;
    return h;
}

void dump_tree(const xa::FTree& t, int* unc, int* inter, int* coupled)
{
    if(unc)
        for(int i = 0; i < t.rank; ++i) q_to(t.unc[i], unc + 2 * i);
    if(inter)
        for(int i = 0; i < t.n_inter(); ++i) q_to(t.inter[i], inter + 2 * i);
    if(coupled) q_to(t.coupled, coupled);
}

}  // namespace

extern "C" {

int xt_rt_create(int device, xt_rt** out)
{
    *out = nullptr;
    auto* r = new xt_rt();
    if(device >= 0) {
        xh_ctx* ctx = nullptr;
        int rc = xh_create(device, &ctx);
        if(rc != XH_OK) {
            g_err = "no GPU available for the tensor runtime";
            delete r;
            return rc;
        }
        r->rt.holder = std::make_shared<xa::CtxHolder>();
        r->rt.holder->ctx = ctx;
        r->rt.ctx = ctx;
    }
    *out = r;
    return XH_OK;
}

int xt_rt_destroy(xt_rt* rt)
{
    if(!rt) return XH_OK;
    if(rt->live > 0) {
        // tensors and CTM objects keep raw pointers to the runtime: refuse rather than dangle
        rt->err = g_err = "xt_rt_destroy: " + std::to_string(rt->live) + " tensor or CTM handle(s) of this runtime are still alive; destroy them first";
        return XH_EINVAL;
    }
    if(rt->rt.ctx) xh_synchronize(rt->rt.ctx);
    // the runtime destructor releases the cached plans; the context itself lives on until the
    // last tensor buffer that uses it is released
    delete rt;
    return XH_OK;
}

const char* xt_last_error(const xt_rt* rt) { return rt ? rt->err.c_str() : g_err.c_str(); }

xh_ctx* xt_rt_ctx(xt_rt* rt) { return rt ? rt->rt.ctx : nullptr; }

int xt_rt_counters(const xt_rt* rt, double* gemm_flops, double* permute_bytes, long* n_gemm, long* n_permute)
{
    if(!rt) return XH_EINVAL;
    if(gemm_flops) *gemm_flops = rt->rt.gemm_flops;
    if(permute_bytes) *permute_bytes = rt->rt.permute_bytes;
    if(n_gemm) *n_gemm = rt->rt.n_gemm;
    if(n_permute) *n_permute = rt->rt.n_permute;
    return XH_OK;
}

int xt_rt_host_stats(const xt_rt* rt, double* permute_plan_s, double* gemm_plan_s, long* n_permute_plans, long* n_gemm_plans)
{
    if(!rt) return XH_EINVAL;
    *permute_plan_s = rt->rt.permute_plan_s;
    *gemm_plan_s = rt->rt.gemm_plan_s;
    *n_permute_plans = rt->rt.n_permute_plans;
    *n_gemm_plans = rt->rt.n_gemm_plans;
    return XH_OK;
}

int xt_rt_host_timers(const xt_rt* rt, double* t, int n)
{
    if(!rt || n < 0) return XH_EINVAL;
    for(int i = 0; i < n && i < 5; ++i) t[i] = rt->rt.host_t[i];
    if(n > 5) t[5] = static_cast<double>(rt->rt.n_symbolic_hits);
    if(n > 6) t[6] = rt->rt.host_t[5];
    if(n > 7) t[7] = rt->rt.host_t[6];
    return XH_OK;
}

int xt_rt_gemm_timing(xt_rt* rt, int enable, int reset, double* kernel_ms, double* flops, long* nlog)
{
    return run(rt, [&] {
        xh_ctx* ctx = rt->rt.ctx;
        if(reset) {
            rt->rt.gemm_log.clear();
            if(ctx && xh_gemm_timing(ctx, -1, 1, nullptr, nullptr, nullptr) != XH_OK) throw xa::Error(XH_EHIP, xh_last_error(ctx));
        }
        if(enable >= 0) rt->rt.time_gemm = enable != 0;  // shape log of the contraction blocks
        double ms = 0.0, fl = 0.0;
        int64_t n = 0;
        if(ctx && xh_gemm_timing(ctx, enable, 0, &ms, &fl, &n) != XH_OK) throw xa::Error(XH_EHIP, xh_last_error(ctx));
        if(kernel_ms) *kernel_ms = ms;
        if(flops) *flops = fl;
        if(nlog) *nlog = static_cast<long>(rt->rt.gemm_log.size());
    });
}

int xt_rt_marker(xt_rt* rt, int id)
{
    return run(rt, [&] {
        if(rt->rt.ctx && xh_trace_marker(rt->rt.ctx, id) != XH_OK) throw xa::Error(XH_EINVAL, xh_last_error(rt->rt.ctx));
    });
}

int xt_rt_gemm_log(const xt_rt* rt, int32_t* mnk, long cap)
{
    if(!rt) return XH_EINVAL;
    const long n = std::min<long>(cap, static_cast<long>(rt->rt.gemm_log.size()));
    for(long i = 0; i < n; ++i)
        for(int j = 0; j < 3; ++j) mnk[3 * i + j] = rt->rt.gemm_log[i][j];
    return XH_OK;
}

int xt_rt_set_max_plans(xt_rt* rt, int64_t max_plans)
{
    return run(rt, [&] {
        if(max_plans < 8) throw xa::Error(XH_EINVAL, "xt_rt_set_max_plans: the bound must be at least 8");
        rt->rt.max_plans = static_cast<std::size_t>(max_plans);
        rt->rt.trim_caches();
    });
}

int xt_rt_set_fuse_permute(xt_rt* rt, int enable, long* n_fused)
{
    return run(rt, [&] {
        if(enable < -1 || enable > 2) throw xa::Error(XH_EINVAL, "xt_rt_set_fuse_permute: enable must be -1, 0, 1 or 2");
        if(enable >= 0) rt->rt.fuse_permute = enable;
        if(n_fused) *n_fused = rt->rt.n_fused;
    });
}

int xt_rt_cache_sizes(const xt_rt* rt, int64_t* permute_plans, int64_t* gemm_plans, int64_t* evicted)
{
    if(!rt) return XH_EINVAL;
    if(permute_plans) *permute_plans = static_cast<int64_t>(rt->rt.permute_cache.size());
    if(gemm_plans) *gemm_plans = static_cast<int64_t>(rt->rt.gemm_cache.size());
    if(evicted) *evicted = rt->rt.n_evicted;
    return XH_OK;
}

int xt_rt_clear_caches(xt_rt* rt)
{
    return run(rt, [&] {
        if(rt->rt.ctx) xh_synchronize(rt->rt.ctx);
        // gather plans first: each refers to a GEMM plan of the cache
        for(auto& [k, v] : rt->rt.gather_cache)
            if(rt->rt.ctx) v->release(rt->rt.ctx);
        rt->rt.gather_cache.clear();
        for(auto& [k, v] : rt->rt.permute_cache)
            if(v->dev) xh_permute_plan_destroy(rt->rt.ctx, v->dev), v->dev = nullptr;
        for(auto& [k, v] : rt->rt.gemm_cache)
            if(v->dev) xh_gemm_plan_destroy(rt->rt.ctx, v->dev), v->dev = nullptr;
        rt->rt.permute_cache.clear();
        rt->rt.sym_permute_cache.clear();
        rt->rt.gemm_cache.clear();
        rt->rt.struct_cache.clear();
    });
}

// ---------------------------------------------------------------- bases

int xt_basis_create(int kind0, int kind1, xt_basis** out)
{
    return run(nullptr, [&] {
        auto* h = new xt_basis();
        h->b = xa::Qbasis(xa::Sym::make(kind0, kind1), 1);
        *out = h;
    });
}

int xt_basis_destroy(xt_basis* b)
{
    delete b;
    return XH_OK;
}

int xt_basis_copy(const xt_basis* b, xt_basis** out)
{
    return run(nullptr, [&] { *out = new xt_basis(*b); });
}

int xt_basis_push_back(xt_basis* b, const int* q, size_t dim)
{
    return run(nullptr, [&] {
        if(b->b.depth != 1) throw xa::Error(XH_EINVAL, "push_back needs a depth-1 basis");
        b->b.push_back(q_from(q), dim);
    });
}

int xt_basis_conj(const xt_basis* b, xt_basis** out)
{
    return run(nullptr, [&] {
        if(b->b.depth != 1) throw xa::Error(XH_EINVAL, "conj needs a depth-1 basis");
        *out = new xt_basis{b->b.conj()};
    });
}

int xt_basis_combine(const xt_basis* a, const xt_basis* b, xt_basis** out)
{
    return run(nullptr, [&] {
        if(b->b.depth != 1) throw xa::Error(XH_EINVAL, "combine: the second basis must have depth 1");
        if(!(a->b.sym == b->b.sym)) throw xa::Error(XH_EINVAL, "combine: symmetry mismatch");
        *out = new xt_basis{a->b.combine(b->b)};
    });
}

int xt_basis_forget_history(const xt_basis* b, xt_basis** out)
{
    return run(nullptr, [&] { *out = new xt_basis{b->b.forget_history()}; });
}

int xt_basis_sort(xt_basis* b)
{
    return run(nullptr, [&] { b->b.sort(); });
}

int xt_basis_sym(const xt_basis* b, int* kind0, int* kind1)
{
    if(!b) return XH_EINVAL;
    *kind0 = b->b.sym.kind[0];
    *kind1 = b->b.sym.kind[1];
    return XH_OK;
}

int xt_basis_info(const xt_basis* b, int* depth, int* nsec, size_t* dim, size_t* full_dim, int* conj)
{
    if(!b) return XH_EINVAL;
    if(depth) *depth = b->b.depth;
    if(nsec) *nsec = static_cast<int>(b->b.data.size());
    if(dim) *dim = b->b.dim();
    if(full_dim) *full_dim = b->b.full_dim();
    if(conj) *conj = b->b.conj_flag ? 1 : 0;
    return XH_OK;
}

int xt_basis_entry(const xt_basis* b, int i, int* q, size_t* num, size_t* dim, int* ntrees)
{
    if(!b || i < 0 || i >= static_cast<int>(b->b.data.size())) return XH_EINVAL;
    const auto& e = b->b.data[i];
    if(q) q_to(e.q, q);
    if(num) *num = e.num;
    if(dim) *dim = e.dim;
    if(ntrees) *ntrees = static_cast<int>(e.trees.size());
    return XH_OK;
}

int xt_basis_tree(const xt_basis* b, int i, int t, int* rank, int* unc, int* inter, int* coupled, uint32_t* dims, int* dual, uint64_t* dim)
{
    if(!b || i < 0 || i >= static_cast<int>(b->b.data.size())) return XH_EINVAL;
    const auto& e = b->b.data[i];
    if(t < 0 || t >= static_cast<int>(e.trees.size())) return XH_EINVAL;
    const xa::FTree& tr = e.trees[t];
    if(rank) *rank = tr.rank;
    dump_tree(tr, unc, inter, coupled);
    for(int k = 0; k < tr.rank; ++k) {
        if(dims) dims[k] = tr.dims[k];
        if(dual) dual[k] = tr.dual[k] ? 1 : 0;
    }
    if(dim) *dim = tr.dim;
    return XH_OK;
}

static int tree_results(const xa::TreeMap& m, int cap, int* n, int* unc, int* inter, int* coupled, double* coeff)
{
    *n = static_cast<int>(m.size());
    if(static_cast<int>(m.size()) > cap) return XH_EINVAL;
    int k = 0;
    for(const auto& [tr, c] : m) {
        dump_tree(tr, unc ? unc + 2 * xa::MAXR * k : nullptr, inter ? inter + 2 * xa::MAXR * k : nullptr, coupled ? coupled + 2 * k : nullptr);
        if(coeff) coeff[k] = c;
        ++k;
    }
    return XH_OK;
}

int xt_tree_permute(const xt_basis* b, int i, int t, const int* p, int np, int cap, int* n, int* unc, int* inter, int* coupled, double* coeff)
{
    int rc = XH_OK;
    int r = run(nullptr, [&] {
        const xa::FTree& tr = b->b.data.at(i).trees.at(t);
        if(np != tr.rank) throw xa::Error(XH_EINVAL, "tree permute: permutation length differs from the tree rank");
        xa::Permutation P(std::vector<int>(p, p + np));
        rc = tree_results(xa::tree_permute(b->b.sym, tr, P), cap, n, unc, inter, coupled, coeff);
    });
    return r != XH_OK ? r : rc;
}

int xt_tree_swap(const xt_basis* b, int i, int t, int pos, int cap, int* n, int* unc, int* inter, int* coupled, double* coeff)
{
    int rc = XH_OK;
    int r = run(nullptr, [&] {
        const xa::FTree& tr = b->b.data.at(i).trees.at(t);
        if(pos < 0 || pos >= tr.rank - 1) throw xa::Error(XH_EINVAL, "tree swap: invalid position");
        rc = tree_results(xa::tree_swap(b->b.sym, tr, pos), cap, n, unc, inter, coupled, coeff);
    });
    return r != XH_OK ? r : rc;
}

double xt_wigner_3j(int a, int b, int c, int d, int e, int f) { return xa::wigner_3j(a, b, c, d, e, f); }
double xt_wigner_6j(int a, int b, int c, int d, int e, int f) { return xa::wigner_6j(a, b, c, d, e, f); }
double xt_wigner_9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) { return xa::wigner_9j(a, b, c, d, e, f, g, h, i); }

double xt_coeff(int kind0, int kind1, int which, const int* q1, const int* q2, const int* q3, const int* q4, const int* q5, const int* q6)
{
    xa::Sym s = xa::Sym::make(kind0, kind1);
    switch(which) {
    case 0: return s.coeff_swap(q_from(q1), q_from(q2), q_from(q3));
    case 1: return s.coeff_recouple(q_from(q1), q_from(q2), q_from(q3), q_from(q4), q_from(q5), q_from(q6));
    case 2: return s.coeff_turn(q_from(q1), q_from(q2), q_from(q3));
    case 3: return s.coeff_FS(q_from(q1));
    case 4: return s.coeff_twist(q_from(q1));
    case 5: return s.triangle(q_from(q1), q_from(q2), q_from(q3)) ? 1.0 : 0.0;
    case 6: return static_cast<double>(s.degeneracy(q_from(q1)));
    default: return 0.0;
    }
}

// ---------------------------------------------------------------- tensors

int xt_tensor_create(xt_rt* rt, int rank, int corank, const xt_basis* const* legs, xt_tensor** out)
{
    return run(rt, [&] {
        if(rank < 0 || corank < 0 || rank > xa::MAXR || corank > xa::MAXR) throw xa::Error(XH_EINVAL, "invalid rank");
        std::vector<xa::Qbasis> d, c;
        xa::Sym s;
        bool have = false;
        for(int i = 0; i < rank + corank; ++i) {
            if(legs[i]->b.depth != 1) throw xa::Error(XH_EINVAL, "tensor legs must be depth-1 bases");
            if(have && !(legs[i]->b.sym == s)) throw xa::Error(XH_EINVAL, "legs with different symmetries");
            s = legs[i]->b.sym;
            have = true;
            (i < rank ? d : c).push_back(legs[i]->b);
        }
        if(!have) throw xa::Error(XH_EINVAL, "a tensor needs at least one leg");
        xa::StructPtr st = xa::cached_structure(&rt->rt, s, d, c);
        *out = wrap(rt, xa::DTensor::zeros(&rt->rt, st));
    });
}

int xt_tensor_destroy(xt_tensor* t)
{
    if(t && t->owner) --t->owner->live;
    delete t;
    return XH_OK;
}

int xt_tensor_copy(const xt_tensor* t, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, t->t); });
}

int xt_tensor_shape(const xt_tensor* t, int* rank, int* corank, int* nsec, size_t* size)
{
    if(!t) return XH_EINVAL;
    if(rank) *rank = t->t.st->rank;
    if(corank) *corank = t->t.st->corank;
    if(nsec) *nsec = static_cast<int>(t->t.st->sector.size());
    if(size) *size = t->t.st->size;
    return XH_OK;
}

int xt_tensor_sector(const xt_tensor* t, int i, int* q, size_t* offset, size_t* rows, size_t* cols)
{
    if(!t || i < 0 || i >= static_cast<int>(t->t.st->sector.size())) return XH_EINVAL;
    if(q) q_to(t->t.st->sector[i], q);
    if(offset) *offset = t->t.st->offsets[i];
    if(rows) *rows = t->t.st->rows[i];
    if(cols) *cols = t->t.st->cols[i];
    return XH_OK;
}

int xt_tensor_leg(const xt_tensor* t, int leg, xt_basis** out)
{
    return run(t ? t->owner : nullptr, [&] {
        const auto& s = *t->t.st;
        if(leg < 0 || leg >= s.rank + s.corank) throw xa::Error(XH_EINVAL, "leg out of range");
        *out = new xt_basis{leg < s.rank ? s.unc_dom[leg] : s.unc_cod[leg - s.rank]};
    });
}

int xt_tensor_coupled(const xt_tensor* t, int side, xt_basis** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = new xt_basis{side == 0 ? t->t.st->dom() : t->t.st->cod()}; });
}

int xt_tensor_upload(xt_tensor* t, const double* host, size_t n)
{
    return run(t ? t->owner : nullptr, [&] { t->t.upload(host, n); });
}

int xt_tensor_download(const xt_tensor* t, double* host, size_t n)
{
    return run(t ? t->owner : nullptr, [&] { t->t.download(host, n); });
}

const double* xt_tensor_device_ptr(const xt_tensor* t) { return t ? t->t.data() : nullptr; }

int xt_tensor_set_random(xt_tensor* t, uint32_t seed, int dist)
{
    return run(t ? t->owner : nullptr, [&] {
        std::mt19937 engine(seed);
        t->t.set_random(engine, dist);
    });
}

int xt_permute(const xt_tensor* t, int shift, const int* p, int np, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::permute(t->t, shift, std::vector<int>(p, p + np))); });
}

int xt_permute_plan_stats(const xt_tensor* t, int shift, const int* p, int np, int64_t* ntasks, int64_t* nsources, int64_t* elems)
{
    return run(t ? t->owner : nullptr, [&] {
        // the runtime's cached path; on a GPU runtime the cached plan gave its host tables to
        // the device plan, so they are rebuilt from the structure
        const std::vector<int> pv(p, p + np);
        const auto held = xa::get_permute_plan(t->t.rt, *t->t.st, shift, pv);  // kept alive while read
        const xa::PermutePlan& cached = held->plan;
        const xa::PermutePlan plan = cached.tasks.empty() && cached.moved_elems > 0 ? xa::make_permute_plan(*t->t.st, shift, pv) : cached;
        if(ntasks) *ntasks = static_cast<int64_t>(plan.tasks.size());
        if(nsources) *nsources = static_cast<int64_t>(plan.sources.size());
        if(elems) *elems = static_cast<int64_t>(plan.moved_elems);
    });
}

int xt_permute_plan_dump(const xt_tensor* t, int shift, const int* p, int np, int64_t* task_dst, int64_t* task_rows, int64_t* task_cols, int32_t* task_nsrc,
                         int64_t* src_off, double* src_coeff)
{
    return run(t ? t->owner : nullptr, [&] {
        // the runtime's cached path (symbolic plans reused across dimension changes); on a GPU
        // runtime the host tables were handed to the device plan and are rebuilt here
        const std::vector<int> pv(p, p + np);
        const auto held = xa::get_permute_plan(t->t.rt, *t->t.st, shift, pv);  // kept alive while read
        const xa::PermutePlan& cached = held->plan;
        const xa::PermutePlan plan = cached.tasks.empty() && cached.moved_elems > 0 ? xa::make_permute_plan(*t->t.st, shift, pv) : cached;
        for(std::size_t i = 0; i < plan.tasks.size(); ++i) {
            task_dst[i] = plan.tasks[i].dst_off;
            task_rows[i] = plan.tasks[i].rows;
            task_cols[i] = plan.tasks[i].cols;
            task_nsrc[i] = plan.tasks[i].nsrc;
        }
        for(std::size_t i = 0; i < plan.sources.size(); ++i) {
            src_off[i] = plan.sources[i].src_off;
            src_coeff[i] = plan.sources[i].coeff;
        }
    });
}

int xt_mul(const xt_tensor* l, const xt_tensor* r, int precision, xt_tensor** out)
{
    return run(l ? l->owner : nullptr, [&] { *out = wrap(l->owner, xa::mul(l->t, r->t, precision)); });
}

int xt_mul_into(const xt_tensor* l, const xt_tensor* r, xt_tensor* out, int precision)
{
    return run(l ? l->owner : nullptr, [&] {
        if(!r || !out) throw xa::Error(XH_EINVAL, "xt_mul_into: null tensor");
        xa::mul_into(l->t, r->t, out->t, precision);
    });
}

int xt_contract(const xt_tensor* a, const int* a1, int n1, const xt_tensor* b, const int* a2, int n2, int rankres, int precision, xt_tensor** out)
{
    return run(a ? a->owner : nullptr, [&] {
        *out = wrap(a->owner, xa::contract(a->t, std::vector<int>(a1, a1 + n1), b->t, std::vector<int>(a2, a2 + n2), rankres, precision));
    });
}

int xt_get_permutations(const int* a1, int n1, int rank1, const int* a2, int n2, int rank2, int rankres, int* p1, int* shift1, int* p2, int* shift2,
                        int* pres, int* shiftres)
{
    return run(nullptr, [&] {
        xa::ContractPerms cp = xa::get_permutations(std::vector<int>(a1, a1 + n1), rank1, std::vector<int>(a2, a2 + n2), rank2, rankres);
        std::copy(cp.p1.begin(), cp.p1.end(), p1);
        std::copy(cp.p2.begin(), cp.p2.end(), p2);
        std::copy(cp.pres.begin(), cp.pres.end(), pres);
        *shift1 = cp.shift1;
        *shift2 = cp.shift2;
        *shiftres = cp.shiftres;
    });
}

int xt_tsvd(const xt_tensor* t, size_t max_keep, double eps_svd, int preserve_multiplets, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag,
            double* trunc_weight, double* entropy)
{
    return run(t ? t->owner : nullptr, [&] {
        xa::SVDResult r = xa::tsvd(t->t, max_keep, eps_svd, preserve_multiplets != 0, trunc_weight ? *trunc_weight : 0.0, entropy ? *entropy : 0.0);
        *U = wrap(t->owner, r.U);
        *S = wrap(t->owner, r.S);
        *Vdag = wrap(t->owner, r.Vdag);
        if(trunc_weight) *trunc_weight = r.trunc_weight;
        if(entropy) *entropy = r.entropy;
    });
}

int xt_spec_create(xt_spec** out)
{
    *out = new xt_spec();
    return XH_OK;
}

int xt_spec_destroy(xt_spec* sp)
{
    delete sp;
    return XH_OK;
}

int xt_spec_size(const xt_spec* sp, int* nsec)
{
    if(!sp || !nsec) return XH_EINVAL;
    *nsec = static_cast<int>(sp->m.size());
    return XH_OK;
}

int xt_spec_entry(const xt_spec* sp, int i, int* q, int* n, double* vals, int cap)
{
    if(!sp || i < 0 || i >= static_cast<int>(sp->m.size())) return XH_EINVAL;
    auto it = sp->m.begin();
    std::advance(it, i);
    if(q) q_to(it->first, q);
    if(n) *n = static_cast<int>(it->second.size());
    if(vals)
        for(int k = 0; k < std::min(cap, static_cast<int>(it->second.size())); ++k) vals[k] = it->second[k];
    return XH_OK;
}

int xt_tsvd_spec(const xt_tensor* t, size_t max_keep, double eps_svd, double* trunc_weight, double* entropy, xt_spec* SVspec, int preserve_multiplets,
                 int return_spec, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag)
{
    return run(t ? t->owner : nullptr, [&] {
        if(return_spec && !SVspec) throw xa::Error(XH_EINVAL, "tSVD: RETURN_SPEC needs an SVspec map");
        xa::SVDResult r = xa::tsvd(t->t, max_keep, eps_svd, preserve_multiplets != 0, trunc_weight ? *trunc_weight : 0.0, entropy ? *entropy : 0.0);
        *U = wrap(t->owner, r.U);
        *S = wrap(t->owner, r.S);
        *Vdag = wrap(t->owner, r.Vdag);
        if(trunc_weight) *trunc_weight = r.trunc_weight;
        if(entropy) *entropy = r.entropy;
        if(return_spec)
            for(const auto& [q, v] : r.spec) SVspec->m.insert({q, v});  // std::map::insert keeps an existing entry, as the reference's SVspec.insert
    });
}

int xt_tensor_save(const xt_tensor* t, const char* path)
{
    return run(t ? t->owner : nullptr, [&] {
        if(!t || !path) throw xa::Error(XH_EINVAL, "xt_tensor_save: null argument");
        xa::save(t->t, path);
    });
}

int xt_tensor_load(xt_rt* rt, const char* path, xt_tensor** out)
{
    return run(rt, [&] {
        if(!rt || !path || !out) throw xa::Error(XH_EINVAL, "xt_tensor_load: null argument");
        *out = wrap(rt, xa::load(&rt->rt, path));
    });
}

int xt_tensor_serialize(const xt_tensor* t, void* buf, size_t cap, size_t* n)
{
    return run(t ? t->owner : nullptr, [&] {
        if(!t || !n) throw xa::Error(XH_EINVAL, "xt_tensor_serialize: null argument");
        const auto b = xa::serialize(t->t);
        *n = b.size();
        if(buf && cap >= b.size()) std::memcpy(buf, b.data(), b.size());
    });
}

int xt_tensor_deserialize(xt_rt* rt, const void* buf, size_t n, xt_tensor** out)
{
    return run(rt, [&] {
        if(!rt || !buf || !out) throw xa::Error(XH_EINVAL, "xt_tensor_deserialize: null argument");
        *out = wrap(rt, xa::deserialize(&rt->rt, static_cast<const char*>(buf), n));
    });
}

int xt_qr_bwd(const xt_tensor* A, const xt_tensor* Q, const xt_tensor* R, const xt_tensor* dQ, const xt_tensor* dR, xt_tensor* dA)
{
    return run(A ? A->owner : nullptr, [&] { xa::qr_bwd(A->t, Q->t, R->t, dQ->t, dR->t, dA->t); });
}

int xt_tsvd_randomized(const xt_tensor* t, size_t max_keep, double eps_svd, const int* qs, const int* ranks, int nsec, int oversample,
                       double rank_factor, int power_iters, int sketch_precision, uint64_t seed, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag,
                       double* trunc_weight, int* safe, double* residual_ratio)
{
    return run(t ? t->owner : nullptr, [&] {
        std::map<xa::Q, int> m;
        for(int i = 0; i < nsec; ++i) m[q_from(qs + 2 * i)] = ranks[i];
        xa::RsvdOptions o;
        o.oversample = oversample;
        o.rank_factor = rank_factor;
        o.power_iters = power_iters;
        o.sketch_precision = sketch_precision;
        xa::RandomizedSVD r = xa::tsvd_randomized(t->t, m, max_keep, eps_svd, o, seed);
        if(r.res.U.st == nullptr) throw xa::Error(XH_EINVAL, "randomized SVD of an empty tensor");
        *U = wrap(t->owner, r.res.U);
        *S = wrap(t->owner, r.res.S);
        *Vdag = wrap(t->owner, r.res.Vdag);
        if(trunc_weight) *trunc_weight = r.res.trunc_weight;
        if(safe) *safe = r.safe ? 1 : 0;
        if(residual_ratio) *residual_ratio = r.max_residual_ratio;
    });
}

int xt_svd_full(const xt_tensor* t, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag)
{
    return run(t ? t->owner : nullptr, [&] {
        xa::FullSVD f = xa::full_svd(t->t);
        *U = wrap(t->owner, f.U);
        *S = wrap(t->owner, f.S);
        *Vdag = wrap(t->owner, f.Vdag);
    });
}

int xt_cutoff(const xt_tensor* t, const xt_tensor* U, const xt_tensor* S, const xt_tensor* Vdag, size_t max_keep, double eps_svd,
              int preserve_multiplets, xt_tensor** Ut, xt_tensor** St, xt_tensor** Vdagt, double* trunc_weight, double* entropy)
{
    return run(t ? t->owner : nullptr, [&] {
        const xa::TensorStructure& ts = *t->t.st;
        const xa::TensorStructure& ss = *S->t.st;
        if(U->t.st->sector != ts.sector || ss.sector != ts.sector || Vdag->t.st->sector != ts.sector)
            throw xa::Error(XH_EINVAL, "cutoff: the factors do not have the sectors of the tensor");
        xa::FullSVD f;
        f.U = U->t;
        f.S = S->t;
        f.Vdag = Vdag->t;
        f.sectors = ts.sector;
        std::vector<double> hs(ss.size);
        if(ss.size) S->t.download(hs.data(), hs.size());
        for(std::size_t i = 0; i < ts.sector.size(); ++i) {
            const std::size_t r = std::min(ts.rows[i], ts.cols[i]);
            if(ss.rows[i] != r || ss.cols[i] != r || U->t.st->cols[i] != r || Vdag->t.st->rows[i] != r)
                throw xa::Error(XH_EINVAL, "cutoff: the factors are not the full SVD of the tensor");
            for(std::size_t k = 0; k < r; ++k) f.all_sv.emplace_back(ts.sector[i], hs[ss.offsets[i] + k + k * r]);
        }
        xa::SVDResult res = xa::cutoff(t->t, f, max_keep, eps_svd, preserve_multiplets != 0, trunc_weight ? *trunc_weight : 0.0, entropy ? *entropy : 0.0);
        *Ut = wrap(t->owner, res.U);
        *St = wrap(t->owner, res.S);
        *Vdagt = wrap(t->owner, res.Vdag);
        if(trunc_weight) *trunc_weight = res.trunc_weight;
        if(entropy) *entropy = res.entropy;
    });
}

int xt_tensor_zeros_like(const xt_tensor* t, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::DTensor::zeros(t->t.rt, t->t.st)); });
}

int xt_permute_adj(const xt_tensor* y, int shift, const int* p, int np, xt_tensor** out)
{
    return run(y ? y->owner : nullptr, [&] { *out = wrap(y->owner, xa::permute_adj(y->t, shift, std::vector<int>(p, p + np))); });
}

int xt_mul_bwd(const xt_tensor* A, const xt_tensor* B, const xt_tensor* dC, xt_tensor* dA, xt_tensor* dB, int precision)
{
    return run(A ? A->owner : nullptr, [&] {
        if(!B || !dC) throw xa::Error(XH_EINVAL, "mul adjoint: null operand");
        xa::mul_bwd(A->t, B->t, dC->t, dA ? &dA->t : nullptr, dB ? &dB->t : nullptr, precision);
    });
}

int xt_svd_bwd(const xt_tensor* U, const xt_tensor* S, const xt_tensor* Vdag, const xt_tensor* dU, const xt_tensor* dS, const xt_tensor* dVdag, xt_tensor* dA)
{
    return run(U ? U->owner : nullptr, [&] {
        if(!S || !Vdag || !dU || !dS || !dVdag || !dA) throw xa::Error(XH_EINVAL, "svd adjoint: null operand");
        xa::svd_bwd(U->t, S->t, Vdag->t, dU->t, dS->t, dVdag->t, dA->t);
    });
}

int xt_cutoff_bwd(const xt_tensor* dUt, const xt_tensor* dSt, const xt_tensor* dVdagt, xt_tensor* dU, xt_tensor* dS, xt_tensor* dVdag)
{
    return run(dUt ? dUt->owner : nullptr, [&] {
        if(!dSt || !dVdagt || !dU || !dS || !dVdag) throw xa::Error(XH_EINVAL, "cutoff adjoint: null operand");
        xa::cutoff_bwd(dUt->t, dSt->t, dVdagt->t, dU->t, dS->t, dVdag->t);
    });
}

int xt_diag_bwd(const xt_tensor* y, const xt_tensor* dy, xt_tensor* dx, int mode)
{
    return run(y ? y->owner : nullptr, [&] {
        if(!dy || !dx) throw xa::Error(XH_EINVAL, "diag adjoint: null operand");
        xa::diag_bwd(y->t, dy->t, dx->t, mode);
    });
}

int xt_trace_bwd(xt_tensor* dx, double g)
{
    return run(dx ? dx->owner : nullptr, [&] { xa::trace_bwd(dx->t, g); });
}

int xt_max_norm_bwd(const xt_tensor* x, xt_tensor* dx, double g)
{
    return run(x ? x->owner : nullptr, [&] {
        if(!dx) throw xa::Error(XH_EINVAL, "maxNorm adjoint: null operand");
        xa::max_norm_bwd(x->t, dx->t, g);
    });
}

int xt_tqr(const xt_tensor* t, int return_lq, xt_tensor** Q, xt_tensor** R)
{
    return run(t ? t->owner : nullptr, [&] {
        auto [q, r] = xa::tqr(t->t, return_lq != 0);
        *Q = wrap(t->owner, q);
        *R = wrap(t->owner, r);
    });
}

int xt_teigh(const xt_tensor* t, xt_tensor** D, xt_tensor** V)
{
    return run(t ? t->owner : nullptr, [&] {
        auto [d, v] = xa::teigh(t->t);
        *D = wrap(t->owner, d);
        *V = wrap(t->owner, v);
    });
}

int xt_adjoint(const xt_tensor* t, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::adjoint(t->t)); });
}

int xt_twist(const xt_tensor* t, int leg, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::twist(t->t, leg)); });
}

int xt_scale(const xt_tensor* t, double alpha, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::scaled(t->t, alpha)); });
}

int xt_axpby(const xt_tensor* x, const xt_tensor* y, double a, double b, xt_tensor** out)
{
    return run(x ? x->owner : nullptr, [&] { *out = wrap(x->owner, xa::axpby(x->t, y->t, a, b)); });
}

int xt_diag_map(const xt_tensor* t, int mode, xt_tensor** out)
{
    return run(t ? t->owner : nullptr, [&] { *out = wrap(t->owner, xa::diag_map(t->t, mode)); });
}

int xt_trace(const xt_tensor* t, double* out)
{
    return run(t ? t->owner : nullptr, [&] { *out = xa::trace(t->t); });
}

int xt_squared_norm(const xt_tensor* t, double* out)
{
    return run(t ? t->owner : nullptr, [&] { *out = xa::squared_norm(t->t); });
}

int xt_max_norm(const xt_tensor* t, double* out)
{
    return run(t ? t->owner : nullptr, [&] { *out = xa::max_norm(t->t); });
}

int xt_truncation_rule(int kind0, int kind1, const int* qs, const double* svs, int64_t n, size_t max_keep, double eps_svd, int preserve_multiplets,
                       int* nsec, int* kept_q, int* kept_n, double* trunc_weight, double* entropy)
{
    return run(nullptr, [&] {
        xa::Sym s = xa::Sym::make(kind0, kind1);
        std::vector<std::pair<xa::Q, double>> all;
        for(int64_t i = 0; i < n; ++i) all.emplace_back(q_from(qs + 2 * i), svs[i]);
        xa::TruncRule r = xa::truncation_rule(s, all, max_keep, eps_svd, preserve_multiplets != 0, *trunc_weight, *entropy);
        int k = 0;
        for(const auto& [q, v] : r.kept) {
            q_to(q, kept_q + 2 * k);
            kept_n[k] = static_cast<int>(v.size());
            ++k;
        }
        *nsec = k;
        *trunc_weight = r.trunc_weight;
        *entropy = r.entropy;
    });
}

// ---------------------------------------------------------------- CTMRG

}  // extern "C"

extern "C" {

int xt_tape_begin(xt_rt* rt)
{
    return run(rt, [&] {
        if(rt->tape) throw xa::Error(XH_EINVAL, "a tape is already recording on this runtime");
        rt->tape = std::make_unique<xa::Tape>();
        rt->tape->active = true;
        rt->rt.tape = rt->tape.get();
    });
}

int xt_tape_pause(xt_rt* rt, int paused)
{
    return run(rt, [&] {
        if(!rt->tape) throw xa::Error(XH_EINVAL, "no tape on this runtime");
        rt->tape->active = paused == 0;
    });
}

int xt_tape_seed(xt_rt* rt, const xt_tensor* y, const xt_tensor* dy, double alpha)
{
    return run(rt, [&] {
        if(!rt->tape) throw xa::Error(XH_EINVAL, "no tape on this runtime");
        rt->tape->accumulate(y->t, dy->t, alpha);
    });
}

int xt_tape_seed_trace(xt_rt* rt, const xt_tensor* y, double g)
{
    return run(rt, [&] {
        if(!rt->tape) throw xa::Error(XH_EINVAL, "no tape on this runtime");
        xa::DTensor d = rt->tape->adjoint_of(y->t);
        xa::trace_bwd(d, g);
    });
}

int xt_tape_backward(xt_rt* rt)
{
    return run(rt, [&] {
        if(!rt->tape) throw xa::Error(XH_EINVAL, "no tape on this runtime");
        rt->tape->backward();
    });
}

int xt_tape_adjoint(xt_rt* rt, const xt_tensor* x, xt_tensor** out)
{
    return run(rt, [&] {
        if(!rt->tape) throw xa::Error(XH_EINVAL, "no tape on this runtime");
        const xa::DTensor* d = rt->tape->find(x->t);
        *out = wrap(rt, d ? *d : xa::DTensor::zeros(&rt->rt, x->t.st));
    });
}

int xt_tape_end(xt_rt* rt)
{
    rt->rt.tape = nullptr;
    rt->tape.reset();
    return XH_OK;
}

int xt_div_trace(const xt_tensor* x, const xt_tensor* s, xt_tensor** out)
{
    return run(x->owner, [&] { *out = wrap(x->owner, xa::div_trace(x->t, s->t)); });
}

}  // extern "C"

struct xt_peps {
    xa::WeightedPEPS p;
    xt_rt* owner = nullptr;
};
struct xt_tprop {
    std::unique_ptr<xa::TimePropagator> tp;
    xt_rt* owner = nullptr;
};

extern "C" {

int xt_tensor_mexp(const xt_tensor* t, double factor, xt_tensor** out)
{
    return run(t->owner, [&] { *out = wrap(t->owner, xa::mexp(t->t, factor)); });
}

int xt_peps_create(xt_rt* rt, const xt_tensor* const* A, size_t D, xt_peps** out)
{
    return run(rt, [&] {
        std::vector<xa::DTensor> a;
        for(int i = 0; i < 4; ++i) a.push_back(A[i]->t);
        auto* p = new xt_peps();
        try {
            p->p = xa::init_weight_tensors(&rt->rt, a, D);
        } catch(...) {
            delete p;
            throw;
        }
        p->owner = rt;
        ++rt->live;
        *out = p;
    });
}

int xt_peps_destroy(xt_peps* p)
{
    if(p && p->owner) --p->owner->live;
    delete p;
    return XH_OK;
}

int xt_peps_tensor(const xt_peps* p, int which, int site, xt_tensor** out)
{
    return run(p->owner, [&] {
        if(site < 0 || site >= 4 || which < 0 || which > 3) throw xa::Error(XH_EINVAL, "peps tensor index out of range");
        if(which == 3) {
            *out = wrap(p->owner, xa::weighted_sites(p->p)[site]);
            return;
        }
        const std::vector<xa::DTensor>* tabs[3] = {&p->p.Gs, &p->p.whs, &p->p.wvs};
        *out = wrap(p->owner, (*tabs[which])[site]);
    });
}

int xt_tprop_create(xt_rt* rt, const xt_tensor* const* Hh, const xt_tensor* const* Hv, double dt, xt_tprop** out)
{
    return run(rt, [&] {
        std::vector<xa::DTensor> h, v;
        for(int i = 0; i < 4; ++i) {
            h.push_back(Hh[i]->t);
            v.push_back(Hv[i]->t);
        }
        auto* t = new xt_tprop();
        try {
            t->tp = std::make_unique<xa::TimePropagator>(&rt->rt, h, v, dt);
        } catch(...) {
            delete t;
            throw;
        }
        t->owner = rt;
        ++rt->live;
        *out = t;
    });
}

int xt_tprop_destroy(xt_tprop* t)
{
    if(t && t->owner) --t->owner->live;
    delete t;
    return XH_OK;
}

int xt_tprop_step(xt_tprop* t, xt_peps* p)
{
    return run(t->owner, [&] { t->tp->t_step(p->p); });
}

int xt_tprop_bond(xt_tprop* t, xt_peps* p, int vertical, int x, int y)
{
    return run(t->owner, [&] {
        if(vertical)
            t->tp->t_step_v(p->p, x, y);
        else
            t->tp->t_step_h(p->p, x, y);
    });
}

int xt_tprop_tensor(const xt_tprop* t, int which, int site, xt_tensor** out)
{
    return run(t->owner, [&] {
        if(site < 0 || site >= 4 || which < 0 || which > 3) throw xa::Error(XH_EINVAL, "propagator tensor index out of range");
        const std::vector<xa::DTensor>* tabs[4] = {&t->tp->Uh, &t->tp->Uv, &t->tp->spectrum_h, &t->tp->spectrum_v};
        const xa::DTensor& r = (*tabs[which])[site];
        if(!r.st) throw xa::Error(XH_EINVAL, "spectrum of a bond that was not updated yet");
        *out = wrap(t->owner, r);
    });
}

}  // extern "C"

struct xt_ctm {
    std::unique_ptr<xa::CTM> ctm;
    xt_rt* owner = nullptr;
};

extern "C" {

int xt_ctm_create(xt_rt* rt, const xt_tensor* const* A, size_t chi, xt_ctm** out)
{
    return run(rt, [&] {
        std::vector<xa::DTensor> a;
        for(int i = 0; i < 4; ++i) a.push_back(A[i]->t);
        auto* c = new xt_ctm();
        c->ctm = std::make_unique<xa::CTM>(&rt->rt, a, chi);
        c->owner = rt;
        ++rt->live;
        *out = c;
    });
}

int xt_ctm_destroy(xt_ctm* c)
{
    if(c && c->owner) --c->owner->live;
    delete c;
    return XH_OK;
}

int xt_ctm_set_options(xt_ctm* c, int precision, int svd_mode, int oversample, int power_iters, int flags)
{
    return run(c->owner, [&] {
        xa::CTMOptions o = c->ctm->opt;
        o.precision = precision;
        o.svd = svd_mode == 1 ? xa::SvdMode::Randomized : xa::SvdMode::Full;
        o.rsvd.oversample = oversample;
        o.rsvd.power_iters = power_iters;
        o.record_spectra = (flags & 1) != 0;
        o.check_finite = (flags & 2) != 0;
        o.verbose = (flags & 4) != 0;
        o.predictor_reuses_spectra = (flags & 8) == 0;
        o.batch_contractions = (flags & 16) != 0;
        c->ctm->set_options(o);
    });
}

int xt_ctm_init(xt_ctm* c)
{
    return run(c->owner, [&] { c->ctm->init_from_A(); });
}

int xt_ctm_grow_all(xt_ctm* c)
{
    return run(c->owner, [&] { c->ctm->grow_all(); });
}

int xt_ctm_compute_rdm(xt_ctm* c)
{
    return run(c->owner, [&] { c->ctm->compute_rdm(); });
}

int xt_ctm_energy(xt_ctm* c, double Jxy, double Jz, double* E)
{
    return run(c->owner, [&] { *E = c->ctm->energy(Jxy, Jz); });
}

int xt_ctm_tensor(xt_ctm* c, int which, int site, xt_tensor** out)
{
    return run(c->owner, [&] {
        if(site < 0 || site >= 4 || which < 0 || which > 9) throw xa::Error(XH_EINVAL, "ctm tensor index out of range");
        xa::CTM& m = *c->ctm;
        std::vector<xa::DTensor>* tabs[10] = {&m.C1, &m.C2, &m.C3, &m.C4, &m.T1, &m.T2, &m.T3, &m.T4, &m.rho_h, &m.rho_v};
        const xa::DTensor& t = (*tabs[which])[site];
        if(!t.st) throw xa::Error(XH_EINVAL, "ctm tensor not computed yet");
        *out = wrap(c->owner, t);
    });
}

int xt_ctm_copy_env(xt_ctm* dst, const xt_ctm* src)
{
    return run(dst->owner, [&] {
        xa::CTM::State st = src->ctm->state();
        if(st.C1.empty() || !st.C1[0].st) throw xa::Error(XH_EINVAL, "copy_env: the source has no environment");
        dst->ctm->set_state(st);
    });
}

int xt_ctm_energy_grad(xt_ctm* c, int steps, double Jxy, double Jz, double* E, xt_tensor** dA)
{
    return run(c->owner, [&] {
        std::vector<xa::DTensor> g;
        *E = c->ctm->energy_gradient(steps, Jxy, Jz, g);
        for(int i = 0; i < 4; ++i) dA[i] = wrap(c->owner, g[i]);
    });
}

int xt_ctm_predictor_stats(const xt_ctm* c, long* from_previous, long* from_sketch)
{
    *from_previous = c->ctm->n_predicted_from_previous;
    *from_sketch = c->ctm->n_predicted_from_sketch;
    return XH_OK;
}

int xt_ctm_svd_ms(const xt_ctm* c, double* ms)
{
    *ms = c->ctm->svd_ms;
    return XH_OK;
}

int xt_ctm_set_rsvd(xt_ctm* c, int oversample, double rank_factor, int power_iters, int sketch_precision)
{
    return run(c->owner, [&] {
        if(oversample < 0 || power_iters < 0 || !(rank_factor >= 0.0)) throw xa::Error(XH_EINVAL, "invalid randomized SVD parameters");
        xa::CTMOptions o = c->ctm->opt;
        o.rsvd.oversample = oversample;
        o.rsvd.rank_factor = rank_factor;
        o.rsvd.power_iters = power_iters;
        o.rsvd.sketch_precision = sketch_precision;
        c->ctm->set_options(o);
    });
}

int xt_ctm_set_predictor(xt_ctm* c, const float* params, int64_t nparams, int width, double margin)
{
    return run(c->owner, [&] {
        if(nparams <= 0) {
            c->ctm->predictor.reset();
            return;
        }
        c->owner->rt.require_device();
        c->ctm->predictor = std::make_shared<xa::RankPredictor>(c->owner->rt.ctx, params, nparams, width, margin);
    });
}

int xt_rank_features(const double* top, int ntop, int dim, double s_glob, double log_cut, float* out)
{
    if(!out || ntop < 0 || (ntop > 0 && !top)) return XH_EINVAL;
    xa::rank_features(top, ntop, dim, s_glob, log_cut, out);
    return XH_OK;
}

int xt_ctm_rsvd_stats(const xt_ctm* c, long* n_rsvd, long* n_fallback, double* max_residual_ratio)
{
    *n_rsvd = c->ctm->n_rsvd;
    *n_fallback = c->ctm->n_fallback;
    *max_residual_ratio = c->ctm->max_residual_ratio;
    return XH_OK;
}

int xt_ctm_num_spectra(const xt_ctm* c, int* n)
{
    *n = static_cast<int>(c->ctm->spectra.size());
    return XH_OK;
}

int xt_ctm_spectrum(const xt_ctm* c, int rec, int* site, int* dir, int* nsec, int* qs, int* rows, int* cols, int* kept, int* nsv, double* sv, int cap)
{
    return run(c->owner, [&] {
        const auto& r = c->ctm->spectra.at(rec);
        *site = r.site;
        *dir = r.dir;
        int k = 0, off = 0;
        for(const auto& [q, v] : r.sv) {
            q_to(q, qs + 2 * k);
            auto sh = r.shape.at(q);
            rows[k] = sh.first;
            cols[k] = sh.second;
            auto it = r.kept.find(q);
            kept[k] = it == r.kept.end() ? 0 : it->second;
            nsv[k] = static_cast<int>(v.size());
            if(off + static_cast<int>(v.size()) > cap) throw xa::Error(XH_EINVAL, "spectrum buffer too small");
            for(double x : v) sv[off++] = x;
            ++k;
        }
        *nsec = k;
    });
}

int xt_ctm_clear_spectra(xt_ctm* c)
{
    c->ctm->spectra.clear();
    return XH_OK;
}

int xt_ctm_set_ranks(xt_ctm* c, int site, int dir, const int* qs, const int* ranks, int nsec)
{
    return run(c->owner, [&] {
        std::map<xa::Q, int> m;
        for(int i = 0; i < nsec; ++i) m[q_from(qs + 2 * i)] = ranks[i];
        c->ctm->last_kept[{site, dir}] = m;
    });
}

}  // extern "C"
