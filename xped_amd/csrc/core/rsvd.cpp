// Randomized truncated block SVD on the GPU (see ctm.hpp).
#include <algorithm>
#include <cmath>
#include <random>

#include "ctm.hpp"

namespace xa {

namespace {

void dchk(xh_ctx* ctx, int rc, const char* what)
{
    if(rc != XH_OK) throw Error(rc, std::string(what) + ": " + xh_last_error(ctx));
}

struct DevBuf {
    xh_ctx* ctx = nullptr;
    double* p = nullptr;
    DevBuf(xh_ctx* c, std::size_t n) : ctx(c)
    {
        if(n) dchk(ctx, xh_alloc(ctx, n * sizeof(double), reinterpret_cast<void**>(&p)), "alloc");
    }
    ~DevBuf()
    {
        if(p) xh_free(ctx, p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

void grouped_gemm(xh_ctx* ctx, const std::vector<xh_gemm_item>& items, int ta, int tb, const double* A, const double* B, double* C, int precision)
{
    if(items.empty()) return;
    xh_gemm_plan* plan = nullptr;
    dchk(ctx, xh_gemm_plan_create(ctx, items.data(), static_cast<int64_t>(items.size()), ta, tb, &plan), "gemm plan");
    int rc = xh_gemm(ctx, plan, A, B, C, 1.0, 0.0, precision);
    xh_gemm_plan_destroy(ctx, plan);
    dchk(ctx, rc, "grouped gemm");
}

}  // namespace

RandomizedSVD tsvd_randomized(const DTensor& t, const std::map<Q, int>& ranks, std::size_t max_keep, double eps_svd, int oversample, int power_iters,
                              int sketch_precision, std::uint32_t seed)
{
    t.rt->require_device();
    xh_ctx* ctx = t.rt->ctx;
    const TensorStructure& s = *t.st;
    struct Sec {
        int i;  // sector index in t
        int64_t m, n, l;
        int64_t om, y, z, b, ub, sb, u;  // offsets into the work buffers
    };
    std::vector<Sec> secs;
    int64_t nOm = 0, nY = 0, nZ = 0, nB = 0, nUb = 0, nSb = 0;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        auto it = ranks.find(s.sector[i]);
        const int k = it == ranks.end() ? 0 : it->second;
        if(k <= 0) continue;
        const int64_t m = static_cast<int64_t>(s.rows[i]), n = static_cast<int64_t>(s.cols[i]);
        const int64_t l = std::min<int64_t>(k + oversample, std::min(m, n));
        if(l <= 0) continue;
        Sec c{static_cast<int>(i), m, n, l, nOm, nY, nZ, nB, nUb, nSb, 0};
        nOm += n * l;
        nY += m * l;
        nZ += n * l;
        nB += l * n;
        nUb += l * l;
        nSb += l;
        secs.push_back(c);
    }
    // Gaussian test matrices from one engine, sector by sector
    std::vector<double> hom(nOm);
    {
        std::mt19937 eng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        for(auto& x : hom) x = nd(eng);
    }
    DevBuf Om(ctx, nOm), Y(ctx, nY), Qy(ctx, nY), Z(ctx, nZ), Qz(ctx, nZ), Rw(ctx, std::max(nUb, int64_t(1))), Bm(ctx, nB), Ub(ctx, nUb), Sb(ctx, nSb),
        Vtb(ctx, nB), Uf(ctx, nY);
    if(nOm) dchk(ctx, xh_upload(ctx, Om.p, hom.data(), nOm * sizeof(double)), "upload sketch");
    const double* A = t.data();
    auto items = [&](auto f) {
        std::vector<xh_gemm_item> v;
        for(const auto& c : secs) v.push_back(f(c));
        return v;
    };
    auto qr = [&](double* src, double* dst, bool rows_m) {
        std::vector<xh_factor_item> it;
        for(const auto& c : secs) {
            const int64_t rr = rows_m ? c.m : c.n;
            const int64_t off = rows_m ? c.y : c.z;
            it.push_back(xh_factor_item{off, off, c.l, c.ub, static_cast<int32_t>(rr), static_cast<int32_t>(c.l)});
        }
        if(!it.empty()) dchk(ctx, xh_block_qr(ctx, it.data(), static_cast<int64_t>(it.size()), src, dst, Rw.p), "block qr");
    };
    const int64_t a0 = 0;
    (void)a0;
    // Y = A Omega
    grouped_gemm(ctx, items([&](const Sec& c) {
                     return xh_gemm_item{static_cast<int64_t>(s.offsets[c.i]), c.om, c.y, static_cast<int32_t>(c.m), static_cast<int32_t>(c.l),
                                         static_cast<int32_t>(c.n)};
                 }),
                 0, 0, A, Om.p, Y.p, sketch_precision);
    for(int p = 0; p < power_iters; ++p) {
        qr(Y.p, Qy.p, true);
        // Z = A^T Q
        grouped_gemm(ctx, items([&](const Sec& c) {
                         return xh_gemm_item{static_cast<int64_t>(s.offsets[c.i]), c.y, c.z, static_cast<int32_t>(c.n), static_cast<int32_t>(c.l),
                                             static_cast<int32_t>(c.m)};
                     }),
                     1, 0, A, Qy.p, Z.p, sketch_precision);
        qr(Z.p, Qz.p, false);
        // Y = A Qz
        grouped_gemm(ctx, items([&](const Sec& c) {
                         return xh_gemm_item{static_cast<int64_t>(s.offsets[c.i]), c.z, c.y, static_cast<int32_t>(c.m), static_cast<int32_t>(c.l),
                                             static_cast<int32_t>(c.n)};
                     }),
                     0, 0, A, Qz.p, Y.p, sketch_precision);
    }
    qr(Y.p, Qy.p, true);
    // B = Q^T A in FP64
    grouped_gemm(ctx, items([&](const Sec& c) {
                     return xh_gemm_item{c.y, static_cast<int64_t>(s.offsets[c.i]), c.b, static_cast<int32_t>(c.l), static_cast<int32_t>(c.n),
                                         static_cast<int32_t>(c.m)};
                 }),
                 1, 0, Qy.p, A, Bm.p, XH_FP64);
    // small SVDs of B (l x n)
    {
        std::vector<xh_factor_item> it;
        for(const auto& c : secs) it.push_back(xh_factor_item{c.b, c.ub, c.sb, c.b, static_cast<int32_t>(c.l), static_cast<int32_t>(c.n)});
        if(!it.empty()) dchk(ctx, xh_block_svd(ctx, it.data(), static_cast<int64_t>(it.size()), Bm.p, Ub.p, Sb.p, Vtb.p), "block svd of the projection");
    }
    // U = Q Ub in FP64
    grouped_gemm(ctx, items([&](const Sec& c) {
                     return xh_gemm_item{c.y, c.ub, c.y, static_cast<int32_t>(c.m), static_cast<int32_t>(c.l), static_cast<int32_t>(c.l)};
                 }),
                 0, 0, Qy.p, Ub.p, Uf.p, XH_FP64);
    // singular values and captured weights
    std::vector<double> sv(nSb);
    if(nSb) dchk(ctx, xh_download(ctx, sv.data(), Sb.p, nSb * sizeof(double)), "download singular values");
    std::vector<double> normA(s.sector.size(), 0.0), normB(secs.size(), 0.0);
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        const std::size_t n = s.rows[i] * s.cols[i];
        if(n) dchk(ctx, xh_dot(ctx, A + s.offsets[i], A + s.offsets[i], n, &normA[i]), "norm");
    }
    for(std::size_t k = 0; k < secs.size(); ++k) {
        double acc = 0.0;
        for(int64_t j = 0; j < secs[k].l; ++j) acc += sv[secs[k].sb + j] * sv[secs[k].sb + j];
        normB[k] = acc;
    }
    std::vector<std::pair<Q, double>> all_sv;
    for(const auto& c : secs)
        for(int64_t j = 0; j < c.l; ++j) all_sv.emplace_back(s.sector[c.i], sv[c.sb + j]);
    RandomizedSVD out;
    if(all_sv.empty()) {
        out.safe = false;
        return out;
    }
    TruncRule rule = truncation_rule(s.sym, all_sv, max_keep, eps_svd, false, 0.0, 0.0);
    double sigma_cut = std::numeric_limits<double>::infinity();
    for(const auto& [q, v] : rule.kept)
        for(double x : v) sigma_cut = std::min(sigma_cut, x);
    // an uncaptured singular value of sector q is at most sqrt(residual_q)
    std::vector<char> sketched(s.sector.size(), 0);
    for(std::size_t k = 0; k < secs.size(); ++k) {
        sketched[secs[k].i] = 1;
        const double res = std::max(0.0, normA[secs[k].i] - normB[k]);
        out.max_residual_ratio = std::max(out.max_residual_ratio, res / (sigma_cut * sigma_cut));
        if(res >= sigma_cut * sigma_cut) out.safe = false;
    }
    for(std::size_t i = 0; i < s.sector.size(); ++i)
        if(!sketched[i]) {
            out.max_residual_ratio = std::max(out.max_residual_ratio, normA[i] / (sigma_cut * sigma_cut));
            if(normA[i] >= sigma_cut * sigma_cut) out.safe = false;
        }
    // assemble U, Sigma, Vdag over the kept sectors (ascending q, as cutoff_matrices does)
    Qbasis truncBasis(s.sym, 1);
    std::vector<Q> order;
    for(const auto& [q, v] : rule.kept) {
        truncBasis.push_back(q, v.size());
        order.push_back(q);
    }
    SVDResult& res = out.res;
    res.trunc_weight = rule.trunc_weight;
    res.entropy = rule.entropy;
    res.spec = rule.kept;
    res.n_states = all_sv.size();
    res.U = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {truncBasis}, order));
    res.S = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {truncBasis}, {truncBasis}, order));
    res.Vdag = DTensor::empty(t.rt, TensorStructure::create_pushed(s.sym, {truncBasis}, s.unc_cod, order));
    std::vector<xh_copy_rect> ru, rv;
    std::vector<double> hs(res.S.size(), 0.0);
    for(std::size_t o = 0; o < order.size(); ++o) {
        const Q& q = order[o];
        const Sec* c = nullptr;
        for(const auto& x : secs)
            if(s.sector[x.i] == q) c = &x;
        const std::size_t nret = rule.kept[q].size();
        ru.push_back(xh_copy_rect{c->y, c->m, static_cast<int64_t>(res.U.st->offsets[o]), c->m, static_cast<int32_t>(c->m), static_cast<int32_t>(nret), 1.0, 0});
        rv.push_back(xh_copy_rect{c->b, c->l, static_cast<int64_t>(res.Vdag.st->offsets[o]), static_cast<int64_t>(nret), static_cast<int32_t>(nret),
                                  static_cast<int32_t>(c->n), 1.0, 0});
        for(std::size_t k = 0; k < nret; ++k) hs[res.S.st->offsets[o] + k + k * nret] = rule.kept[q][k];
    }
    dchk(ctx, xh_copy_rects(ctx, Uf.p, res.U.data(), ru.data(), static_cast<int64_t>(ru.size())), "truncate U");
    dchk(ctx, xh_copy_rects(ctx, Vtb.p, res.Vdag.data(), rv.data(), static_cast<int64_t>(rv.size())), "truncate Vdag");
    if(res.S.size()) res.S.upload(hs.data(), hs.size());
    return out;
}

}  // namespace xa
