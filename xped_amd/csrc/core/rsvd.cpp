// Randomized truncated block SVD on the GPU.
//
// Per sector block A_q (m x n) with target rank k_q the range finder works on l_q columns,
// l_q = min(ceil(rank_factor * k_q) + oversample, m, n):
//   Y = A Omega (Omega Gaussian, n x l)            Q = orth(Y)
//   repeat power_iters times: Z = A^T Q, Q_z = orth(Z), Y = A Q_z, Q = orth(Y)
//   B^T = A^T Q (n x l)                            Q_b = orth(B^T),  R = Q_b^T B^T (l x l)
//   R = U_r S V_r^T (Jacobi)                       A ~ (Q V_r) S (Q_b U_r)^T
// orth is shifted CholeskyQR3 (Fukaya et al., SIAM J. Sci. Comput. 42 (2020) A477): three passes
// of Gram GEMM, Cholesky-inverse of the l x l Gram block and a GEMM that applies it (one shifted
// pass inside the subspace iteration, where only the span matters).  Every step
// is one grouped launch over all sectors.  The range-finder GEMMs (Y, Z) run in the sketch
// precision; B^T, R, the orthonormalisation and the final products run in FP64.
//
// The reference's truncation rule (Tensor.cpp:523-633) then runs over the computed singular
// values.  A sector whose sketch is narrower than the block is trusted when its smallest Ritz
// value sigma_l(B), the estimate of sigma_l(A) >= sigma_{l+1}(A) that subspace iteration
// delivers, stays a safety factor below the smallest kept singular value: otherwise a singular
// value the sketch did not see could belong to the kept set and the caller takes the full SVD.
// (The Frobenius norm of the residual is no usable bound here: the slowly decaying tails of
// environment spectra carry far more weight than sigma_cut^2 while sigma_{l+1} is well below
// sigma_cut.)
#include <algorithm>
#include <cmath>
#include <limits>

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
    DevBuf(xh_ctx* c, int64_t n) : ctx(c)
    {
        if(n > 0) dchk(ctx, xh_alloc(ctx, static_cast<size_t>(n) * sizeof(double), reinterpret_cast<void**>(&p)), "alloc");
    }
    ~DevBuf()
    {
        if(p) xh_free(ctx, p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

struct Plan {
    xh_ctx* ctx = nullptr;
    xh_gemm_plan* p = nullptr;
    Plan(xh_ctx* c, const std::vector<xh_gemm_item>& items, int ta, int tb) : ctx(c)
    {
        if(!items.empty()) dchk(ctx, xh_gemm_plan_create(ctx, items.data(), static_cast<int64_t>(items.size()), ta, tb, &p), "gemm plan");
    }
    ~Plan()
    {
        if(p) xh_gemm_plan_destroy(ctx, p);
    }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    void run(const double* A, const double* B, double* C, int precision, double alpha = 1.0, double beta = 0.0) const
    {
        if(p) dchk(ctx, xh_gemm(ctx, p, A, B, C, alpha, beta, precision), "grouped gemm");
    }
};

struct Sec {
    int t;                   // input tensor of the batch
    int i;                   // sector index in the input tensor
    int64_t a, m, n, l;      // block offset and extents
    int64_t ym, zn, g, s;    // offsets: m x l panels, n x l panels, l x l blocks, l vectors
};

}  // namespace

RandomizedSVD tsvd_randomized(const DTensor& t, const std::map<Q, int>& ranks, std::size_t max_keep, double eps_svd, const RsvdOptions& o,
                              std::uint64_t seed)
{
    return std::move(tsvd_randomized_batch({&t}, {ranks}, max_keep, eps_svd, o, {seed})[0]);
}

std::vector<RandomizedSVD> tsvd_randomized_batch(const std::vector<const DTensor*>& ts, const std::vector<std::map<Q, int>>& ranks_of,
                                                 std::size_t max_keep, double eps_svd, const RsvdOptions& o, const std::vector<std::uint64_t>& seeds)
{
    const int nt = static_cast<int>(ts.size());
    if(nt == 0) return {};
    if(ranks_of.size() != ts.size() || seeds.size() != ts.size()) throw Error(XH_EINVAL, "randomized SVD batch: one rank map and one seed per tensor");
    Runtime* rt = ts[0]->rt;
    rt->require_device();
    for(const DTensor* t : ts)
        if(t->rt != rt) throw Error(XH_EINVAL, "randomized SVD batch: tensors of different runtimes");
    xh_ctx* ctx = rt->ctx;
    // the grouped launches address every tensor's blocks from the lowest data address
    const double* A = nullptr;
    for(const DTensor* t : ts)
        if(t->size() && (!A || t->data() < A)) A = t->data();
    std::vector<Sec> secs;
    std::vector<int64_t> zbeg(nt + 1, 0);
    int64_t nY = 0, nZ = 0, nG = 0, nS = 0;
    for(int ti = 0; ti < nt; ++ti) {
        const TensorStructure& s = *ts[ti]->st;
        zbeg[ti] = nZ;
        if(!ts[ti]->size()) continue;
        const int64_t tbase = static_cast<int64_t>((reinterpret_cast<std::uintptr_t>(ts[ti]->data()) - reinterpret_cast<std::uintptr_t>(A)) / sizeof(double));
        for(std::size_t i = 0; i < s.sector.size(); ++i) {
            const int64_t m = static_cast<int64_t>(s.rows[i]), n = static_cast<int64_t>(s.cols[i]);
            if(m == 0 || n == 0) continue;
            auto it = ranks_of[ti].find(s.sector[i]);
            const int k = it == ranks_of[ti].end() ? 0 : std::max(0, it->second);
            const int64_t want = static_cast<int64_t>(std::ceil(o.rank_factor * k)) + o.oversample;
            const int64_t l = std::min<int64_t>(std::max<int64_t>(want, 1), std::min(m, n));
            secs.push_back(Sec{ti, static_cast<int>(i), tbase + static_cast<int64_t>(s.offsets[i]), m, n, l, nY, nZ, nG, nS});
            nY += m * l;
            nZ += n * l;
            nG += l * l;
            nS += l;
        }
    }
    zbeg[nt] = nZ;
    std::vector<RandomizedSVD> outs(nt);
    if(secs.empty()) {
        for(auto& out : outs) out.safe = false;
        return outs;
    }
    auto items = [&](auto f) {
        std::vector<xh_gemm_item> v;
        v.reserve(secs.size());
        for(const auto& c : secs) v.push_back(f(c));
        return v;
    };
    auto I32 = [](int64_t x) { return static_cast<int32_t>(x); };
    // grouped GEMM plans (offsets are shared by all panels of one kind)
    const Plan p_sketch(ctx, items([&](const Sec& c) { return xh_gemm_item{c.a, c.zn, c.ym, I32(c.m), I32(c.l), I32(c.n)}; }), 0, 0);  // Y = A Om
    const Plan p_atq(ctx, items([&](const Sec& c) { return xh_gemm_item{c.a, c.ym, c.zn, I32(c.n), I32(c.l), I32(c.m)}; }), 1, 0);    // Z = A^T Q
    const Plan p_aqz(ctx, items([&](const Sec& c) { return xh_gemm_item{c.a, c.zn, c.ym, I32(c.m), I32(c.l), I32(c.n)}; }), 0, 0);    // Y = A Qz
    const Plan p_gram_m(ctx, items([&](const Sec& c) { return xh_gemm_item{c.ym, c.ym, c.g, I32(c.l), I32(c.l), I32(c.m)}; }), 1, 0);
    const Plan p_apply_m(ctx, items([&](const Sec& c) { return xh_gemm_item{c.ym, c.g, c.ym, I32(c.m), I32(c.l), I32(c.l)}; }), 0, 0);
    const Plan p_gram_n(ctx, items([&](const Sec& c) { return xh_gemm_item{c.zn, c.zn, c.g, I32(c.l), I32(c.l), I32(c.n)}; }), 1, 0);
    const Plan p_apply_n(ctx, items([&](const Sec& c) { return xh_gemm_item{c.zn, c.g, c.zn, I32(c.n), I32(c.l), I32(c.l)}; }), 0, 0);
    const Plan p_u(ctx, items([&](const Sec& c) { return xh_gemm_item{c.ym, c.g, c.ym, I32(c.m), I32(c.l), I32(c.l)}; }), 0, 1);      // U = Q V_r
    std::vector<xh_factor_item> chol_m, chol_n, svd_r;
    std::vector<double> shift_m, shift_n, zero(secs.size(), 0.0);
    for(const auto& c : secs) {
        chol_m.push_back(xh_factor_item{c.g, c.g, 0, 0, I32(c.l), I32(c.l)});
        svd_r.push_back(xh_factor_item{c.g, c.g, c.s, c.g, I32(c.l), I32(c.l)});
        const double eps = std::numeric_limits<double>::epsilon();
        const double ll = static_cast<double>(c.l);
        shift_m.push_back(11.0 * (static_cast<double>(c.m) * ll + ll * (ll + 1.0)) * eps);
        shift_n.push_back(11.0 * (static_cast<double>(c.n) * ll + ll * (ll + 1.0)) * eps);
    }
    chol_n = chol_m;
    DevBuf Y0(ctx, nY), Y1(ctx, nY), Y2(ctx, nY), Z0(ctx, nZ), Z1(ctx, nZ), Z2(ctx, nZ), G(ctx, nG), X(ctx, nG), Ur(ctx, nG), Vtr(ctx, nG),
        Sr(ctx, nS);
    // shifted CholeskyQR3 of the panels in src; the result lands in b1 (src is left untouched)
    auto orth = [&](const double* src, double* b1, double* b2, const Plan& gram, const Plan& apply, const std::vector<xh_factor_item>& chol,
                    const std::vector<double>& shift, int passes) {
        const double* in = src;
        double* outs[3] = {b1, b2, b1};
        for(int pass = 0; pass < passes; ++pass) {
            gram.run(in, in, G.p, XH_FP64);
            dchk(ctx, xh_block_chol_inv(ctx, chol.data(), static_cast<int64_t>(chol.size()), G.p, X.p, pass == 0 ? shift.data() : zero.data()),
                 "cholesky inverse");
            apply.run(in, X.p, outs[pass], XH_FP64);
            in = outs[pass];
        }
        return b1;
    };
    // range finder
    for(int ti = 0; ti < nt; ++ti)
        if(zbeg[ti + 1] > zbeg[ti])
            dchk(ctx, xh_fill_gaussian(ctx, Z0.p + zbeg[ti], static_cast<size_t>(zbeg[ti + 1] - zbeg[ti]), seeds[ti]), "gaussian test matrix");
    p_sketch.run(A, Z0.p, Y0.p, o.sketch_precision);
    // inside the subspace iteration one shifted CholeskyQR pass keeps the panels well conditioned;
    // the basis that B is projected on gets all three passes
    double* Qm = orth(Y0.p, Y1.p, Y2.p, p_gram_m, p_apply_m, chol_m, shift_m, o.power_iters > 0 ? 1 : 3);
    for(int p = 0; p < o.power_iters; ++p) {
        p_atq.run(A, Qm, Z0.p, o.sketch_precision);
        double* Qn = orth(Z0.p, Z1.p, Z2.p, p_gram_n, p_apply_n, chol_n, shift_n, 1);
        p_aqz.run(A, Qn, Y0.p, o.sketch_precision);
        Qm = orth(Y0.p, Y1.p, Y2.p, p_gram_m, p_apply_m, chol_m, shift_m, p + 1 < o.power_iters ? 1 : 3);
    }
    // B^T = A^T Q in FP64
    p_atq.run(A, Qm, Z0.p, XH_FP64);
    // QR of B^T and the small SVD
    double* Qb = orth(Z0.p, Z1.p, Z2.p, p_gram_n, p_apply_n, chol_n, shift_n, 3);
    p_gram_n.run(Qb, Z0.p, G.p, XH_FP64);  // R = Q_b^T B^T
    dchk(ctx, xh_block_svd(ctx, svd_r.data(), static_cast<int64_t>(svd_r.size()), G.p, Ur.p, Sr.p, Vtr.p), "block svd of R");
    p_u.run(Qm, Vtr.p, Y0.p, XH_FP64);     // U = Q V_r   (m x l)
    p_apply_n.run(Qb, Ur.p, Z0.p, XH_FP64);  // V = Q_b U_r (n x l)
    // singular values and the captured weights
    std::vector<double> sv(nS);
    dchk(ctx, xh_download(ctx, sv.data(), Sr.p, static_cast<size_t>(nS) * sizeof(double)), "download singular values");
    for(int ti = 0; ti < nt; ++ti) {
        const TensorStructure& s = *ts[ti]->st;
        RandomizedSVD& out = outs[ti];
        if(!ts[ti]->size()) {
            out.safe = false;
            continue;
        }
        std::vector<std::pair<Q, double>> all_sv;
        for(const auto& c : secs)
            if(c.t == ti)
                for(int64_t j = 0; j < c.l; ++j) {
                    all_sv.emplace_back(s.sector[c.i], sv[c.s + j]);
                    out.ritz[s.sector[c.i]].push_back(sv[c.s + j]);
                }
        TruncRule rule = truncation_rule(s.sym, all_sv, max_keep, eps_svd, false, 0.0, 0.0);
        double sigma_cut = std::numeric_limits<double>::infinity();
        for(const auto& [q, v] : rule.kept)
            for(double x : v) sigma_cut = std::min(sigma_cut, x);
        for(const auto& c : secs) {
            if(c.t != ti || c.l >= std::min(c.m, c.n)) continue;  // the sketch spans the whole block
            const double ratio = sv[c.s + c.l - 1] / sigma_cut;
            out.max_residual_ratio = std::max(out.max_residual_ratio, ratio);
            if(!(o.safety * ratio < 1.0)) out.safe = false;
        }
        // U, Sigma, Vdag over the kept sectors (ascending q, as cutoff_matrices builds them)
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
        res.U = DTensor::empty(rt, TensorStructure::create_pushed(s.sym, s.unc_dom, {truncBasis}, order));
        res.S = DTensor::empty(rt, TensorStructure::create_pushed(s.sym, {truncBasis}, {truncBasis}, order));
        res.Vdag = DTensor::empty(rt, TensorStructure::create_pushed(s.sym, {truncBasis}, s.unc_cod, order));
        std::vector<xh_copy_rect> ru, rv;
        std::vector<double> hs(res.S.size(), 0.0);
        for(std::size_t o2 = 0; o2 < order.size(); ++o2) {
            const Q& q = order[o2];
            const Sec* c = nullptr;
            for(const auto& x : secs)
                if(x.t == ti && s.sector[x.i] == q) c = &x;
            if(!c) throw Error(XH_EINVAL, "randomized SVD: kept sector without a sketch");
            const int64_t nret = static_cast<int64_t>(rule.kept[q].size());
            ru.push_back(xh_copy_rect{c->ym, c->m, static_cast<int64_t>(res.U.st->offsets[o2]), c->m, I32(c->m), I32(nret), 1.0, 0});
            rv.push_back(xh_copy_rect{c->zn, c->n, static_cast<int64_t>(res.Vdag.st->offsets[o2]), nret, I32(nret), I32(c->n), 1.0, 1});
            for(int64_t k = 0; k < nret; ++k) hs[res.S.st->offsets[o2] + k + k * nret] = rule.kept[q][k];
        }
        dchk(ctx, xh_copy_rects(ctx, Y0.p, res.U.data(), ru.data(), static_cast<int64_t>(ru.size())), "truncate U");
        dchk(ctx, xh_copy_rects(ctx, Z0.p, res.Vdag.data(), rv.data(), static_cast<int64_t>(rv.size())), "truncate Vdag");
        if(res.S.size()) res.S.upload(hs.data(), hs.size());
    }
    return outs;
}

}  // namespace xa
