// CTMRG driver on the GPU tensor layer (see ctm.hpp for the reference functions it follows).
//
// The contraction index lists below restate the reference's network (CTM.cpp, TRank = 2,
// ALL_OUT_LEGS = false).  Twists are kept where the reference applies them; they are the
// identity for bosonic symmetries.
#include "ctm.hpp"

#include "ad.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace xa {

namespace {
enum Leg { L_LEFT = 0, L_TOP = 1, L_RIGHT = 2, L_BOTTOM = 3, L_PHYS = 4 };
enum Dir { D_LEFT = 0, D_RIGHT = 1, D_TOP = 2, D_BOTTOM = 3 };
enum Corner { UPPER_LEFT = 0, LOWER_LEFT = 1, UPPER_RIGHT = 2, LOWER_RIGHT = 3 };

DTensor tw(const DTensor& t, std::initializer_list<int> legs)
{
    DTensor out = t;
    for(int l : legs) out = twist(out, l);
    return out;
}
}  // namespace

CTM::CTM(Runtime* r, const std::vector<DTensor>& A, std::size_t chi) : rt(r)
{
    if(A.size() != Lx * Ly) throw Error(XH_EINVAL, "CTM: expected 4 site tensors for the 2x2 unit cell");
    opt.chi = chi;
    As = A;
    // iPEPS::updateAdags (iPEPS.cpp:200-214): Adag = A^dagger permuted to (left, top, phys | right, bottom)
    for(const auto& a : As) Adags.push_back(permute(adjoint(a), 0, {3, 4, 2, 0, 1}));
    const int n = Lx * Ly;
    C1.resize(n);
    C2.resize(n);
    C3.resize(n);
    C4.resize(n);
    T1.resize(n);
    T2.resize(n);
    T3.resize(n);
    T4.resize(n);
    rho_h.resize(n);
    rho_v.resize(n);
}

DTensor CTM::ct(const DTensor& a, std::vector<int> a1, const DTensor& b, std::vector<int> a2, int rank) const
{
    return contract(a, a1, b, a2, rank, opt.precision);
}

DTensor CTM::identity(const std::vector<Qbasis>& dom, const std::vector<Qbasis>& cod) const
{
    return xa::identity(rt, dom, cod);
}

Qbasis CTM::ket(int x, int y, int leg) const
{
    const TensorStructure& s = *A(x, y).st;
    switch(leg) {
    case L_LEFT: return s.unc_dom[0];
    case L_TOP: return s.unc_dom[1];
    case L_RIGHT: return s.unc_cod[0];
    case L_BOTTOM: return s.unc_cod[1];
    default: return s.unc_cod[2];
    }
}

Qbasis CTM::bra(int x, int y, int leg) const
{
    const TensorStructure& s = *Adag(x, y).st;
    switch(leg) {
    case L_LEFT: return s.unc_dom[0];
    case L_TOP: return s.unc_dom[1];
    case L_RIGHT: return s.unc_cod[0];
    case L_BOTTOM: return s.unc_cod[1];
    default: return s.unc_dom[2];
    }
}

DTensor CTM::normalized(const DTensor& t) const
{
    // the reference divides by maxNorm unchecked (CTM.cpp:1372, 1383, ...).  Here the norm stays on
    // the device; a vanished or non-finite one is counted there and check_normalized() reports it
    // (before the next projector batch and at the end of every move) instead of propagating inf/NaN
    if(!rt->ctx || !t.size()) {
        const double m = max_norm(t);
        if(!(m > 0.0) || !std::isfinite(m)) throw Error(XH_EINVAL, "CTM: environment tensor with max norm " + std::to_string(m) + " cannot be normalized");
        return scaled(t, 1.0 / m);
    }
    DTensor out = DTensor::empty(t.rt, t.st);
    if(xh_normalize_max_abs(rt->ctx, t.data(), out.data(), t.size()) != XH_OK) throw Error(XH_EHIP, std::string("CTM normalize: ") + xh_last_error(rt->ctx));
    if(Tape* tp = recording(t))
        tp->backs.push_back([t, out](Tape& tape) {
            // t * (1 / t.maxNorm()) (ADTensor.hpp:450-476 and the scalar product): with m = |t_k| the
            // largest modulus, dt += dy / m and dt_k += -<dy, t> / m^2 * sign(t_k)
            const DTensor* d = tape.find(out);
            if(!d) return;
            const DTensor dd = *d;
            const double m = max_norm(t);
            double dot = 0.0;
            if(xh_dot(t.rt->ctx, dd.data(), t.data(), t.size(), &dot) != XH_OK) throw Error(XH_EHIP, std::string("CTM normalize adjoint: ") + xh_last_error(t.rt->ctx));
            tape.accumulate(t, dd, 1.0 / m);
            DTensor dt = tape.adjoint_of(t);
            max_norm_bwd(t, dt, -dot / (m * m));
        });
    return out;
}

void CTM::check_normalized() const
{
    if(!rt->ctx) return;
    int64_t count = 0;
    double first = 0.0;
    if(xh_normalize_status(rt->ctx, &count, &first) != XH_OK) throw Error(XH_EHIP, std::string("CTM normalize status: ") + xh_last_error(rt->ctx));
    if(count) throw Error(XH_EINVAL, "CTM: environment tensor with max norm " + std::to_string(first) + " cannot be normalized");
}

void CTM::init_from_A()
{
    for(int x = 0; x < Lx; ++x)
        for(int y = 0; y < Ly; ++y) {
            const int p = idx(x, y);
            const DTensor& a = A(x, y);
            const DTensor& ad = Adag(x, y);
            auto fused = [&](int leg) { return ket(x, y, leg).combine(bra(x, y, leg)).forget_history(); };
            DTensor fuse_ll = identity({fused(L_LEFT)}, {ket(x, y, L_LEFT), bra(x, y, L_LEFT)});
            DTensor fuse_uu = identity({fused(L_TOP)}, {ket(x, y, L_TOP), bra(x, y, L_TOP)});
            DTensor fuse_rr = identity({ket(x, y, L_RIGHT), bra(x, y, L_RIGHT)}, {fused(L_RIGHT)});
            DTensor fuse_dd = identity({ket(x, y, L_BOTTOM), bra(x, y, L_BOTTOM)}, {fused(L_BOTTOM)});
            C1[p] = ct(ct(ct(a, {1, 2, -1, -2, 3}, tw(ad, {3, 4}), {1, 2, 3, -3, -4}, 2), {1, -1, 2, -2}, fuse_rr, {1, 2, -3}, 2), {1, 2, -2}, fuse_dd,
                       {1, 2, -1}, 0);
            C2[p] = ct(ct(ct(a, {-1, 1, 2, -2, 3}, tw(ad, {4}), {-3, 1, 3, 2, -4}, 2), {1, -1, 2, -2}, tw(fuse_ll, {1, 2}), {-3, 1, 2}, 2), {1, 2, -1},
                       fuse_dd, {1, 2, -2}, 1);
            C3[p] = ct(ct(ct(a, {-1, -2, 1, 2, 3}, ad, {-3, -4, 3, 1, 2}, 2), {1, -1, 2, -2}, tw(fuse_ll, {1, 2}), {-3, 1, 2}, 2), {1, 2, -2},
                       tw(fuse_uu, {1, 2}), {-1, 1, 2}, 2);
            C4[p] = ct(ct(ct(a, {1, -1, -2, 2, 3}, tw(ad, {3}), {1, -3, 3, -4, 2}, 2), {1, -1, 2, -2}, tw(fuse_uu, {1, 2}), {-3, 1, 2}, 2), {1, 2, -1},
                       fuse_rr, {1, 2, -2}, 1);
            T1[p] = ct(ct(ct(a, {-1, 1, -2, -3, 2}, tw(ad, {3, 4}), {-4, 1, 2, -5, -6}, 3), {1, -1, -2, 2, -3, -4}, tw(fuse_ll, {1, 2}), {-5, 1, 2}, 4),
                       {1, -3, 2, -4, -1}, fuse_rr, {1, 2, -2}, 1);
            T2[p] = ct(ct(ct(a, {-1, -2, 1, -3, 2}, tw(ad, {4}), {-4, -5, 2, 1, -6}, 3), {-1, 1, -2, -3, 2, -4}, tw(fuse_uu, {1, 2}), {-5, 1, 2}, 4),
                       {-1, 1, -2, 2, -3}, fuse_dd, {1, 2, -4}, 3);
            T3[p] = ct(ct(ct(a, {-1, -2, -3, 1, 2}, tw(ad, {3}), {-4, -5, 2, -6, 1}, 3), {1, -1, -2, 2, -3, -4}, tw(fuse_ll, {1, 2}), {-5, 1, 2}, 4),
                       {-1, 1, -2, 2, -3}, fuse_rr, {1, 2, -4}, 3);
            T4[p] = ct(ct(ct(a, {1, -1, -2, -3, 2}, tw(ad, {3, 4}), {1, -4, 2, -5, -6}, 3), {1, -1, -2, 2, -3, -4}, tw(fuse_uu, {1, 2}), {-5, 1, 2}, 4),
                       {-3, 1, -4, 2, -1}, fuse_dd, {1, 2, -2}, 1);
        }
}

ContractJob CTM::corner_stage(int x, int y, int corner, int stage, const DTensor& prev) const
{
    // contractCorner (CTM.cpp:1643-1760): four contractions per corner; stage k contracts the
    // result of stage k - 1 (prev) with the next environment or site tensor
    switch(corner) {
    case UPPER_LEFT:
        switch(stage) {
        case 0: return {C1[idx(x - 1, y - 1)], {-1, 1}, T1[idx(x, y - 1)], {1, -2, -3, -4}, 1};                // C1T1
        case 1: return {T4[idx(x - 1, y)], {1, -1, -2, -3}, tw(prev, {0}), {1, -4, -5, -6}, 3};                 // T4C1T1
        case 2: return {prev, {-1, 1, -2, -3, 2, -4}, A(x, y), {1, 2, -5, -6, -7}, 4};                         // T4C1T1A
        default: return {prev, {-1, 1, -4, 2, -5, -2, 3}, tw(Adag(x, y), {3, 4}), {1, 2, 3, -6, -3}, 3};
        }
    case LOWER_LEFT:
        switch(stage) {
        case 0: return {C4[idx(x - 1, y + 1)], {-1, 1}, T3[idx(x, y + 1)], {-2, -3, 1, -4}, 1};                // C4T3
        case 1: return {T4[idx(x - 1, y)], {-1, 1, -2, -3}, prev, {1, -4, -5, -6}, 3};                        // T4C4T3
        case 2: return {prev, {-1, 1, -2, 2, -3, -4}, tw(A(x, y), {3}), {1, -5, -6, 2, -7}, 4};                // T4C4T3A
        default: return {prev, {-4, 1, 2, -1, -5, -2, 3}, tw(Adag(x, y), {3}), {1, -6, 3, -3, 2}, 3};
        }
    case UPPER_RIGHT:
        switch(stage) {
        case 0: return {T1[idx(x, y - 1)], {-1, 1, -2, -3}, C2[idx(x + 1, y - 1)], {1, -4}, 3};                // T1C2
        case 1: return {prev, {-1, -2, -3, 1}, T2[idx(x + 1, y)], {-4, -5, 1, -6}, 3};                        // T1C2T2
        case 2: return {prev, {-1, 1, -2, 2, -3, -4}, tw(A(x, y), {2}), {-5, 1, 2, -6, -7}, 4};                // T1C2T2A
        default: return {prev, {-1, 1, 2, -4, -2, -5, 3}, tw(Adag(x, y), {4}), {-3, 1, 3, 2, -6}, 3};
        }
    default:
        switch(stage) {
        case 0: return {tw(C3[idx(x + 1, y + 1)], {1}), {-1, 1}, T3[idx(x, y + 1)], {-2, -3, -4, 1}, 1};       // C3T3
        case 1: return {T2[idx(x + 1, y)], {-1, -2, -3, 1}, prev, {1, -4, -5, -6}, 3};                        // T2C3T3
        case 2: return {prev, {1, -1, -2, 2, -3, -4}, tw(A(x, y), {2, 3}), {-5, -6, 1, 2, -7}, 4};             // T2C3T3A
        default: return {prev, {1, -1, 2, -4, -5, -2, 3}, Adag(x, y), {-6, -3, 3, 1, 2}, 3};
        }
    }
}

std::vector<DTensor> CTM::batched(const std::vector<ContractJob>& jobs) const
{
    if(opt.batch_contractions) return contract_batch(jobs, opt.precision);
    std::vector<DTensor> out;
    for(const ContractJob& j : jobs) out.push_back(contract(j.a, j.a1, j.b, j.a2, j.rank, opt.precision));
    return out;
}

std::vector<DTensor> CTM::contract_corners(const std::vector<std::array<int, 3>>& req) const
{
    // the corners of several sites in lockstep: each stage's contractions share one GEMM launch
    std::vector<DTensor> cur(req.size());
    for(int stage = 0; stage < 4; ++stage) {
        std::vector<ContractJob> jobs;
        for(std::size_t i = 0; i < req.size(); ++i) jobs.push_back(corner_stage(req[i][0], req[i][1], req[i][2], stage, cur[i]));
        cur = batched(jobs);
    }
    return cur;
}

DTensor CTM::contract_corner(int x, int y, int corner) const
{
    return contract_corners({{x, y, corner}})[0];
}

void CTM::keep_spectrum(int site, int dir, const FullSVD& f)
{
    // the leading singular values per sector (all_sv lists each sector's values in descending order)
    std::map<Q, std::vector<double>> m;
    for(const auto& [q, sv] : f.all_sv) {
        auto& v = m[q];
        if(v.size() < static_cast<std::size_t>(RANK_TOP)) v.push_back(sv);
    }
    last_spectra[{site, dir}] = std::move(m);
}

std::pair<DTensor, DTensor> CTM::decompose(const DTensor& Ta, const DTensor& Tb, int site, int dir)
{
    return std::move(decompose_batch({{Ta, Tb}}, {site}, dir)[0]);
}

std::vector<std::pair<DTensor, DTensor>> CTM::decompose_batch(const std::vector<std::pair<DTensor, DTensor>>& in, const std::vector<int>& sites,
                                                              int dir)
{
    // PEPSContractions.cpp:16-43, for every pair: prod = T1 * T2, tSVD(chi, 1e-10, .,
    // PRESERVE_MULTIPLETS = false), P1 = S^-1/2 U^dagger T1, P2 = T2 V S^-1/2.  The pairs of one
    // move are independent; their randomized SVDs share grouped launches.
    const std::size_t nb = in.size();
    if(sites.size() != nb) throw Error(XH_EINVAL, "CTM: one site per projector pair");
    check_normalized();
    std::vector<DTensor> prods(nb);
    for(std::size_t b = 0; b < nb; ++b) {
        prods[b] = mul(in[b].first, in[b].second, opt.precision);
        const double m = max_norm(prods[b]);
        if(!std::isfinite(m))
            throw Error(XH_EINVAL, "CTM: non-finite projector product at site " + std::to_string(sites[b]) + ", direction " + std::to_string(dir));
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<SVDResult> rs(nb);
    auto record = [&](std::size_t b, const SVDResult& r, const FullSVD* f) {
        const DTensor& prod = prods[b];
        std::map<Q, int> kept;
        for(const auto& [q, v] : r.spec) kept[q] = static_cast<int>(v.size());
        last_kept[{sites[b], dir}] = kept;
        if(opt.record_spectra) {
            // a randomized SVD records the kept singular values (it computes no others exactly)
            SpectrumRecord rec;
            rec.site = sites[b];
            rec.dir = dir;
            if(f)
                for(const auto& [q, sv] : f->all_sv) rec.sv[q].push_back(sv);
            else
                for(const auto& [q, v] : r.spec) rec.sv[q] = v;
            rec.kept = kept;
            for(std::size_t i = 0; i < prod.st->sector.size(); ++i)
                rec.shape[prod.st->sector[i]] = {static_cast<int>(prod.st->rows[i]), static_cast<int>(prod.st->cols[i])};
            spectra.push_back(std::move(rec));
        }
    };
    if(opt.svd == SvdMode::Full) {
        for(std::size_t b = 0; b < nb; ++b) {
            FullSVD f = full_svd(prods[b]);
            rs[b] = cutoff(prods[b], f, opt.chi, 1.e-10, false, 0.0, 0.0);
            record(b, rs[b], &f);
            keep_spectrum(sites[b], dir, f);
        }
    } else {
        std::vector<std::map<Q, int>> ranks(nb);
        std::vector<std::uint64_t> seeds(nb);
        std::vector<const DTensor*> ptrs(nb);
        if(!opt.rank_policy && predictor) {
            // the network's features: the previous decomposition's leading singular values of the
            // same (site, direction) when they cover every sector, else a fresh sketch (the sketches
            // of the move's remaining decompositions in one batch)
            std::vector<std::uint64_t> pseeds;
            std::vector<const DTensor*> pp;
            std::vector<std::size_t> need;
            for(std::size_t b = 0; b < nb; ++b) {
                const TensorStructure& st = *prods[b].st;
                auto it = opt.predictor_reuses_spectra ? last_spectra.find({sites[b], dir}) : last_spectra.end();
                bool covered = it != last_spectra.end();
                for(std::size_t i = 0; covered && i < st.sector.size(); ++i) {
                    const std::size_t dim = std::min(st.rows[i], st.cols[i]);
                    if(dim == 0) continue;
                    auto sv = it->second.find(st.sector[i]);
                    covered = sv != it->second.end() && sv->second.size() >= std::min<std::size_t>(dim, RANK_TOP);
                }
                if(covered) {
                    ranks[b] = predictor->ranks_from_spectra(st, it->second, opt.chi);
                    ++n_predicted_from_previous;
                } else {
                    const std::uint64_t nr = static_cast<std::uint64_t>(n_rsvd) + b;
                    pseeds.push_back(0xfeedull + 64ull * nr + 4ull * sites[b] + dir);
                    pp.push_back(&prods[b]);
                    need.push_back(b);
                }
            }
            if(!need.empty()) {
                std::vector<std::map<Q, int>> r = predictor->predict_batch(pp, opt.chi, pseeds);
                for(std::size_t j = 0; j < need.size(); ++j) ranks[need[j]] = std::move(r[j]);
                n_predicted_from_sketch += static_cast<long>(need.size());
            }
        }
        for(std::size_t b = 0; b < nb; ++b) {
            const int site = sites[b];
            const std::uint64_t nr = static_cast<std::uint64_t>(n_rsvd) + b;
            if(opt.rank_policy) {
                ranks[b] = opt.rank_policy(prods[b], site, dir);
            } else if(!predictor) {  // (with a predictor the ranks were set above)
                auto it = last_kept.find({site, dir});
                if(it == last_kept.end()) throw Error(XH_EINVAL, "randomized SVD needs target ranks (no rank policy and no previous full SVD)");
                ranks[b] = it->second;
            }
            seeds[b] = 0x5eedull + 64ull * nr + 4ull * site + dir;
            ptrs[b] = &prods[b];
        }
        std::vector<RandomizedSVD> rsv = tsvd_randomized_batch(ptrs, ranks, opt.chi, 1.e-10, opt.rsvd, seeds);
        n_rsvd += static_cast<decltype(n_rsvd)>(nb);
        for(std::size_t b = 0; b < nb; ++b) {
            max_residual_ratio = std::max(max_residual_ratio, rsv[b].max_residual_ratio);
            if(rsv[b].safe) {
                rs[b] = std::move(rsv[b].res);
                last_spectra[{sites[b], dir}] = std::move(rsv[b].ritz);
            } else {
                // the full FP64 tSVD is the fallback and the judge
                ++n_fallback;
                FullSVD f = full_svd(prods[b]);
                rs[b] = cutoff(prods[b], f, opt.chi, 1.e-10, false, 0.0, 0.0);
                keep_spectrum(sites[b], dir, f);
            }
            record(b, rs[b], nullptr);
        }
    }
    if(rt->ctx) xh_synchronize(rt->ctx);
    const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    svd_ms += dt;
    std::vector<std::pair<DTensor, DTensor>> out(nb);
    for(std::size_t b = 0; b < nb; ++b) {
        const SVDResult& r = rs[b];
        if(opt.verbose) {
            std::size_t kept = 0;
            for(const auto& [q, v] : r.spec) kept += v.size();
            std::fprintf(stderr, "[ctm] site %d dir %d: %zu sectors, %zu kept, svd %.1f ms (batch of %zu)\n", sites[b], dir, prods[b].st->sector.size(), kept,
                         dt, nb);
            std::fflush(stderr);
        }
        const DTensor& Ta = in[b].first;
        const DTensor& Tb = in[b].second;
        DTensor isqrtS = diag_map(r.S, 2);  // S.diag_sqrt().diag_inv()
        DTensor P1 = mul(mul(isqrtS, adjoint(r.U), opt.precision), Ta, opt.precision);
        DTensor P2 = mul(mul(Tb, adjoint(r.Vdag), opt.precision), isqrtS, opt.precision);
        if(opt.check_finite) {
            const std::pair<const char*, const DTensor*> parts[] = {{"U", &r.U}, {"S", &r.S}, {"Vdag", &r.Vdag}, {"P1", &P1}, {"P2", &P2}};
            for(const auto& [name, t] : parts)
                if(!std::isfinite(max_norm(*t)))
                    throw Error(XH_EINVAL, std::string("CTM: non-finite ") + name + " at site " + std::to_string(sites[b]) + ", direction " +
                                               std::to_string(dir));
        }
        out[b] = {std::move(P1), std::move(P2)};
    }
    return out;
}

std::vector<std::pair<DTensor, DTensor>> CTM::projector_inputs_all(const std::vector<std::pair<int, int>>& xy, int dir)
{
    // get_projectors (CTM.cpp:1185-1337): the two corners of each site, all sites' corners batched
    std::vector<std::array<int, 3>> req;
    for(const auto& [x, y] : xy) switch(dir) {
        case D_LEFT:
            req.push_back({x, y, UPPER_LEFT});
            req.push_back({x, y + 1, LOWER_LEFT});
            break;
        case D_RIGHT:
            req.push_back({x, y, UPPER_RIGHT});
            req.push_back({x, y + 1, LOWER_RIGHT});
            break;
        case D_TOP:
            req.push_back({x, y, UPPER_LEFT});
            req.push_back({x + 1, y, UPPER_RIGHT});
            break;
        default:
            req.push_back({x, y, LOWER_LEFT});
            req.push_back({x + 1, y, LOWER_RIGHT});
            break;
        }
    const std::vector<DTensor> Q = contract_corners(req);
    std::vector<std::pair<DTensor, DTensor>> out;
    for(std::size_t s = 0; s < xy.size(); ++s) {
        const DTensor &a = Q[2 * s], &b = Q[2 * s + 1];
        switch(dir) {
        case D_LEFT: out.emplace_back(tw(b, {3, 4, 5}), a); break;  // (Q4, Q1)
        case D_RIGHT: out.emplace_back(a, b); break;                // (Q2, Q3)
        case D_TOP: out.emplace_back(a, b); break;                  // (Q1, Q2)
        default: out.emplace_back(b, tw(a, {0, 1, 2})); break;      // (Q3, Q4)
        }
    }
    return out;
}

std::pair<DTensor, DTensor> CTM::projector_inputs(int x, int y, int dir)
{
    return projector_inputs_all({{x, y}}, dir)[0];
}

std::pair<DTensor, DTensor> CTM::get_projectors(int x, int y, int dir)
{
    auto [Ta, Tb] = projector_inputs(x, y, dir);
    return decompose(Ta, Tb, idx(x, y), dir);
}

void CTM::get_projectors_all(const std::vector<std::pair<int, int>>& xy, int dir, std::vector<DTensor>& P1, std::vector<DTensor>& P2)
{
    std::vector<std::pair<DTensor, DTensor>> in = projector_inputs_all(xy, dir);
    std::vector<int> sites;
    for(const auto& [x, y] : xy) sites.push_back(idx(x, y));
    std::vector<std::pair<DTensor, DTensor>> out = decompose_batch(in, sites, dir);
    for(std::size_t b = 0; b < out.size(); ++b) {
        P1[sites[b]] = std::move(out[b].first);
        P2[sites[b]] = std::move(out[b].second);
    }
}

void CTM::run_chains(std::vector<Chain>& chains)
{
    std::size_t len = 0;
    for(const Chain& c : chains) len = std::max(len, c.stages.size());
    std::vector<DTensor> cur(chains.size());
    for(std::size_t k = 0; k < len; ++k) {
        std::vector<ContractJob> jobs;
        std::vector<std::size_t> who;
        for(std::size_t i = 0; i < chains.size(); ++i)
            if(k < chains[i].stages.size()) {
                jobs.push_back(chains[i].stages[k](cur[i]));
                who.push_back(i);
            }
        std::vector<DTensor> res = batched(jobs);
        for(std::size_t j = 0; j < who.size(); ++j) cur[who[j]] = std::move(res[j]);
    }
    for(std::size_t i = 0; i < chains.size(); ++i) *chains[i].out = normalized(cur[i]);
}

void CTM::left_move()
{
    // CTM.cpp:924-988; projector and new-tensor tables are indexed like the unit cell
    for(int x = 0; x < Lx; ++x) {
        std::vector<DTensor> P1(Lx * Ly), P2(Lx * Ly), C1n(Lx * Ly), T4n(Lx * Ly), C4n(Lx * Ly);
        {
            std::vector<std::pair<int, int>> xy;
            for(int y = 0; y < Ly; ++y) xy.emplace_back(x, y);
            get_projectors_all(xy, D_LEFT, P1, P2);
        }
        std::vector<Chain> chains;
        for(int y = 0; y < Ly; ++y) {
            // renormalize_left (CTM.cpp:1339-1416)
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{C1[idx(x - 1, y - 1)], {-1, 1}, T1[idx(x, y - 1)], {1, -4, -2, -3}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{tw(P1[idx(x, y - 1)], {1, 2, 3}), {-1, 1, 2, 3}, p, {1, 2, 3, -2}, 0}; }},
                             &C1n[idx(x, y - 1)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{C4[idx(x - 1, y + 1)], {-2, 1}, T3[idx(x, y + 1)], {-3, -4, 1, -1}, 1}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-2, 1, 2, 3}, tw(P2[idx(x, y)], {0, 1, 2}), {1, 2, 3, -1}, 1}; }},
                             &C4n[idx(x, y + 1)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{tw(P1[idx(x, y)], {1, 2, 3}), {-1, 1, -2, -3}, T4[idx(x - 1, y)], {-4, 1, -5, -6}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 2, -2, -3, 1, -4}, A(x, y), {1, -5, -6, 2, -7}, 4}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 2, -4, 1, -5, -2, 3}, tw(Adag(x, y), {3, 4}), {1, -6, 3, -3, 2}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-2, -3, -4, 1, 2, 3}, tw(P2[idx(x, y - 1)], {0, 1, 2}), {1, 2, 3, -1}, 1}; }},
                             &T4n[idx(x, y)]});
        }
        run_chains(chains);
        for(int y = 0; y < Ly; ++y) {
            C1[idx(x, y)] = C1n[idx(x, y)];
            T4[idx(x, y)] = T4n[idx(x, y)];
            C4[idx(x, y)] = C4n[idx(x, y)];
        }
    }
}

void CTM::right_move()
{
    // CTM.cpp:990-1051: columns are visited from x = Lx (== 0) downwards; every column once
    for(int xs = Lx; xs > 0; --xs) {
        const int x = xs % Lx;
        std::vector<DTensor> P1(Lx * Ly), P2(Lx * Ly), C2n(Lx * Ly), T2n(Lx * Ly), C3n(Lx * Ly);
        {
            std::vector<std::pair<int, int>> xy;
            for(int y = 0; y < Ly; ++y) xy.emplace_back(x, y);
            get_projectors_all(xy, D_RIGHT, P1, P2);
        }
        std::vector<Chain> chains;
        for(int y = 0; y < Ly; ++y) {
            // renormalize_right (CTM.cpp:1418-1488)
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{T1[idx(x, y - 1)], {-1, 1, -3, -4}, C2[idx(x + 1, y - 1)], {1, -2}, 1}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, 2, 3}, P2[idx(x, y - 1)], {1, 2, 3, -2}, 1}; }},
                             &C2n[idx(x, y - 1)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{tw(C3[idx(x + 1, y + 1)], {1}), {-1, 1}, T3[idx(x, y + 1)], {-2, -3, -4, 1}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{P1[idx(x, y)], {-1, 1, 2, 3}, p, {1, 2, 3, -2}, 2}; }},
                             &C3n[idx(x, y + 1)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{P1[idx(x, y - 1)], {-1, 1, -2, -3}, T2[idx(x + 1, y)], {-4, -5, 1, -6}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, -2, 2, -3, -4}, tw(A(x, y), {2}), {-5, 1, 2, -6, -7}, 4}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, 2, -4, -2, -5, 3}, tw(Adag(x, y), {4}), {-3, 1, 3, 2, -6}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-3, -1, -2, 1, 2, 3}, P2[idx(x, y)], {1, 2, 3, -4}, 3}; }},
                             &T2n[idx(x, y)]});
        }
        run_chains(chains);
        for(int y = 0; y < Ly; ++y) {
            C2[idx(x, y)] = C2n[idx(x, y)];
            T2[idx(x, y)] = T2n[idx(x, y)];
            C3[idx(x, y)] = C3n[idx(x, y)];
        }
    }
}

void CTM::top_move()
{
    // CTM.cpp:1053-1117
    for(int y = 0; y < Ly; ++y) {
        std::vector<DTensor> P1(Lx * Ly), P2(Lx * Ly), C1n(Lx * Ly), T1n(Lx * Ly), C2n(Lx * Ly);
        {
            std::vector<std::pair<int, int>> xy;
            for(int x = 0; x < Lx; ++x) xy.emplace_back(x, y);
            get_projectors_all(xy, D_TOP, P1, P2);
        }
        std::vector<Chain> chains;
        for(int x = 0; x < Lx; ++x) {
            // renormalize_top (CTM.cpp:1490-1562)
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{T4[idx(x - 1, y)], {1, -1, -3, -4}, tw(C1[idx(x - 1, y - 1)], {0}), {1, -2}, 1}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, 2, 3}, P2[idx(x - 1, y)], {1, 2, 3, -2}, 0}; }},
                             &C1n[idx(x - 1, y)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{C2[idx(x + 1, y - 1)], {-1, 1}, T2[idx(x + 1, y)], {-2, -3, 1, -4}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{P1[idx(x, y)], {-1, 1, 2, 3}, p, {1, 2, 3, -2}, 1}; }},
                             &C2n[idx(x + 1, y)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{P1[idx(x - 1, y)], {-1, 1, -2, -3}, T1[idx(x, y - 1)], {1, -4, -5, -6}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, -2, -3, 2, -4}, A(x, y), {1, 2, -5, -6, -7}, 4}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, -4, 2, -5, -2, 3}, tw(Adag(x, y), {3, 4}), {1, 2, 3, -6, -3}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, -3, -4, 1, 2, 3}, P2[idx(x, y)], {1, 2, 3, -2}, 1}; }},
                             &T1n[idx(x, y)]});
        }
        run_chains(chains);
        for(int x = 0; x < Lx; ++x) {
            C1[idx(x, y)] = C1n[idx(x, y)];
            T1[idx(x, y)] = T1n[idx(x, y)];
            C2[idx(x, y)] = C2n[idx(x, y)];
        }
    }
}

void CTM::bottom_move()
{
    // CTM.cpp:1119-1183: rows are visited from y = Ly (== 0) downwards; every row once
    for(int ys = Ly; ys > 0; --ys) {
        const int y = ys % Ly;
        std::vector<DTensor> P1(Lx * Ly), P2(Lx * Ly), C4n(Lx * Ly), T3n(Lx * Ly), C3n(Lx * Ly);
        {
            std::vector<std::pair<int, int>> xy;
            for(int x = 0; x < Lx; ++x) xy.emplace_back(x, y);
            get_projectors_all(xy, D_BOTTOM, P1, P2);
        }
        std::vector<Chain> chains;
        for(int x = 0; x < Lx; ++x) {
            // renormalize_bottom (CTM.cpp:1564-1641)
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{T4[idx(x - 1, y)], {-4, 1, -2, -3}, C4[idx(x - 1, y + 1)], {1, -1}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{tw(P1[idx(x - 1, y)], {1, 2, 3}), {-2, 1, 2, 3}, p, {1, 2, 3, -1}, 1}; }},
                             &C4n[idx(x - 1, y)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{T2[idx(x + 1, y)], {-3, -4, -1, 1}, C3[idx(x + 1, y + 1)], {1, -2}, 1}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, 2, 3}, tw(P2[idx(x, y)], {0, 1, 2}), {1, 2, 3, -2}, 2}; }},
                             &C3n[idx(x + 1, y)]});
            chains.push_back({{[&, x, y](const DTensor&) { return ContractJob{tw(P1[idx(x, y)], {1, 2, 3}), {-1, 1, -2, -3}, tw(T3[idx(x, y + 1)], {0}), {-4, -5, -6, 1}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, -2, 2, -3, -4}, A(x, y), {-5, -6, 1, 2, -7}, 4}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-1, 1, 2, -4, -5, -2, 3}, tw(Adag(x, y), {3}), {-6, -3, 3, 1, 2}, 3}; },
                              [&, x, y](const DTensor& p) { return ContractJob{p, {-4, -1, -2, 1, 2, 3}, tw(P2[idx(x - 1, y)], {0, 1, 2}), {1, 2, 3, -3}, 3}; }},
                             &T3n[idx(x, y)]});
        }
        run_chains(chains);
        for(int x = 0; x < Lx; ++x) {
            C4[idx(x, y)] = C4n[idx(x, y)];
            T3[idx(x, y)] = T3n[idx(x, y)];
            C3[idx(x, y)] = C3n[idx(x, y)];
        }
    }
}

void CTM::grow_all()
{
    left_move();
    check_normalized();
    right_move();
    check_normalized();
    top_move();
    check_normalized();
    bottom_move();
    check_normalized();
}

void CTM::compute_rdm()
{
    auto normalize_rho = [&](DTensor rho) {
        // rho * (1 / rho.contract(Id2.twist(0).twist(1)).trace())
        DTensor Id2 = identity(rho.st->unc_cod, rho.st->unc_dom);
        return div_trace(rho, ct(rho, {1, 2, 3, 4}, tw(Id2, {0, 1}), {3, 4, 1, 2}, 0));
    };
    for(int x = 0; x < Lx; ++x)
        for(int y = 0; y < Ly; ++y) {
            // computeRDM_h (CTM.cpp:528-640, TRank = 2)
            DTensor C1T1 = ct(C1[idx(x - 1, y - 1)], {-1, 1}, T1[idx(x, y - 1)], {1, -2, -3, -4}, 1);
            DTensor T4C1T1 = ct(T4[idx(x - 1, y)], {1, -1, -2, -3}, tw(C1T1, {0}), {1, -4, -5, -6}, 3);
            DTensor T4C1T1A = ct(T4C1T1, {-1, 1, -2, -3, 2, -4}, A(x, y), {1, 2, -5, -6, -7}, 4);
            DTensor C4T3 = ct(C4[idx(x - 1, y + 1)], {-1, 1}, T3[idx(x, y + 1)], {-2, -3, 1, -4}, 1);
            DTensor C4T3Ad = ct(C4T3, {-1, -2, 1, -3}, tw(Adag(x, y), {3}), {-4, -5, -6, -7, 1}, 3);
            DTensor left_half = ct(T4C1T1A, {1, 2, -1, 3, -2, 4, -3}, C4T3Ad, {1, 4, -4, 2, 3, -5, -6}, 3);
            DTensor T1C2 = ct(T1[idx(x + 1, y - 1)], {-1, 1, -2, -3}, C2[idx(x + 2, y - 1)], {1, -4}, 3);
            DTensor T1C2T2 = ct(T1C2, {-1, -2, -3, 1}, T2[idx(x + 2, y)], {-4, -5, 1, -6}, 3);
            DTensor AT1C2T2 = ct(tw(A(x + 1, y), {1}), {-1, 1, 2, -2, -3}, T1C2T2, {-4, 1, -5, 2, -6, -7}, 3);
            DTensor T3C3 = ct(T3[idx(x + 1, y + 1)], {-1, -2, -3, 1}, C3[idx(x + 2, y + 1)], {-4, 1}, 3);
            DTensor T3C3Ad = ct(T3C3, {-1, 1, -2, -3}, Adag(x + 1, y), {-4, -5, -6, -7, 1}, 3);
            DTensor right_half = ct(AT1C2T2, {-1, 1, -2, -3, 2, 3, 4}, T3C3Ad, {1, -4, 4, -5, 2, -6, 3}, 3);
            rho_h[idx(x, y)] = normalize_rho(ct(left_half, {1, 2, -3, 3, -1, 4}, right_half, {2, -4, 1, 3, 4, -2}, 2));
        }
    for(int x = 0; x < Lx; ++x)
        for(int y = 0; y < Ly; ++y) {
            // computeRDM_v (CTM.cpp:706-800, TRank = 2)
            DTensor C1T1 = ct(C1[idx(x - 1, y - 1)], {-1, 1}, T1[idx(x, y - 1)], {1, -2, -3, -4}, 1);
            DTensor T4C1T1 = ct(T4[idx(x - 1, y)], {1, -1, -2, -3}, tw(C1T1, {0}), {1, -4, -5, -6}, 3);
            DTensor T4C1T1A = ct(T4C1T1, {-1, 1, -2, -3, 2, -4}, A(x, y), {1, 2, -5, -6, -7}, 4);
            DTensor C2T2 = ct(C2[idx(x + 1, y - 1)], {-1, 1}, T2[idx(x + 1, y)], {-2, -3, 1, -4}, 1);
            DTensor AdC2T2 = ct(tw(Adag(x, y), {3, 4}), {-1, -2, -3, 1, -4}, C2T2, {-5, -6, 1, -7}, 4);
            DTensor upper_half = ct(T4C1T1A, {-1, 1, 2, 3, 4, -2, -3}, AdC2T2, {1, 3, -4, -5, 2, 4, -6}, 3);
            DTensor C4T3 = ct(C4[idx(x - 1, y + 2)], {-1, 1}, T3[idx(x, y + 2)], {-2, -3, 1, -4}, 1);
            DTensor T4C4T3 = ct(T4[idx(x - 1, y + 1)], {-1, 1, -2, -3}, C4T3, {1, -4, -5, -6}, 3);
            DTensor AT4C4T3 = ct(tw(A(x, y + 1), {0}), {1, -1, -2, 2, -3}, T4C4T3, {-4, 1, -5, 2, -6, -7}, 3);
            DTensor T2C3 = ct(T2[idx(x + 1, y + 1)], {-1, -2, -3, 1}, C3[idx(x + 1, y + 2)], {1, -4}, 3);
            DTensor T2C3Ad = ct(T2C3, {-1, 1, -2, -3}, Adag(x, y + 1), {-4, -5, -6, 1, -7}, 3);
            DTensor lower_half = ct(AT4C4T3, {-1, 1, -2, -3, 2, 3, 4}, T2C3Ad, {1, -4, 4, 2, -5, -6, 3}, 3);
            rho_v[idx(x, y)] = normalize_rho(ct(upper_half, {1, 2, -3, -1, 3, 4}, lower_half, {2, -4, 1, 4, 3, -2}, 2));
        }
}

std::vector<DTensor> CTM::bond_terms(double Jxy, double Jz) const
{
    // Heisenberg gate Jz Sz.Sz + Jxy/2 (S+S- + S-S+) for spin 1/2 (Heisenberg.hpp:47-53), built on
    // the bases of the reduced density matrix; labels of a U(1) leg are 2 Sz (negated on a
    // conjugated leg).  <H_bond> = rho.contract<{1,2,3,4},{3,4,1,2},0>(gate).trace()
    // (LinearAlgebra.cpp:218-246); E = (sum_h + sum_v) / number of sites (CTMSolver.cpp:60-62).
    auto bond = [&](const DTensor& rho) {
        const TensorStructure& rs = *rho.st;
        StructPtr gs = cached_structure(rt, rs.sym, rs.unc_cod, rs.unc_dom);
        std::vector<double> h(gs->size, 0.0);
        for(std::size_t i = 0; i < gs->sector.size(); ++i) {
            const Q& q = gs->sector[i];
            const auto& dtrees = gs->dom().tree(q);
            const auto& ctrees = gs->cod().tree(q);
            std::size_t r0 = 0;
            for(const auto& td : dtrees) {
                std::size_t c0 = 0;
                for(const auto& tc : ctrees) {
                    if(td.dim == 1 && tc.dim == 1) {
                        auto sz = [&](const Qbasis& b, const Q& lab) { return 0.5 * (b.conj_flag ? -lab.v[0] : lab.v[0]); };
                        const double c1 = sz(gs->unc_dom[0], td.unc[0]), c2 = sz(gs->unc_dom[1], td.unc[1]);
                        const double a1 = sz(gs->unc_cod[0], tc.unc[0]), a2 = sz(gs->unc_cod[1], tc.unc[1]);
                        double v = 0.0;
                        if(c1 == a1 && c2 == a2) v = Jz * a1 * a2;
                        if(c1 == a1 + 1 && c2 == a2 - 1) v = 0.5 * Jxy;  // S+ S-
                        if(c1 == a1 - 1 && c2 == a2 + 1) v = 0.5 * Jxy;  // S- S+
                        h[gs->offsets[i] + r0 + c0 * gs->rows[i]] = v;
                    }
                    c0 += tc.dim;
                }
                r0 += td.dim;
            }
        }
        DTensor gate = DTensor::empty(rt, gs);
        if(gs->size) gate.upload(h.data(), h.size());
        return ct(tw(rho, {0, 1}), {1, 2, 3, 4}, gate, {3, 4, 1, 2}, 0);
    };
    std::vector<DTensor> out;
    for(int i = 0; i < Lx * Ly; ++i) out.push_back(bond(rho_h[i]));
    for(int i = 0; i < Lx * Ly; ++i) out.push_back(bond(rho_v[i]));
    return out;
}

double CTM::energy(double Jxy, double Jz) const
{
    double e = 0.0;
    for(const DTensor& t : bond_terms(Jxy, Jz)) e += trace(t);
    return e / (Lx * Ly);
}

double CTM::step()
{
    grow_all();
    compute_rdm();
    return energy();
}

CTM::State CTM::state() const
{
    return State{C1, C2, C3, C4, T1, T2, T3, T4};
}

void CTM::set_state(const State& st)
{
    C1 = st.C1;
    C2 = st.C2;
    C3 = st.C3;
    C4 = st.C4;
    T1 = st.T1;
    T2 = st.T2;
    T3 = st.T3;
    T4 = st.T4;
}

double CTM::energy_gradient(int steps, double Jxy, double Jz, std::vector<DTensor>& dA)
{
    // CTMSolver::solve with AD (CTMSolver.cpp:80-137): the environment reached so far is a constant;
    // `steps` further grow_all moves, computeRDM and the energy are differentiated with respect to
    // the site tensors.  Each move is a checkpoint (CTM.cpp:386-420 with CP): the forward pass keeps
    // only the environment before every move, and the reverse pass recomputes one move at a time
    // under the tape, seeds the adjoints of its outputs with those carried back from the next move,
    // and runs its reverse passes.
    if(steps < 1) throw Error(XH_EINVAL, "energy_gradient: at least one tracked step");
    if(opt.svd != SvdMode::Full) throw Error(XH_EINVAL, "energy_gradient: the reverse pass needs the full block SVD");
    if(rt->tape) throw Error(XH_EINVAL, "energy_gradient: the runtime is already recording");
    if(C1.empty() || !C1[0].st) throw Error(XH_EINVAL, "energy_gradient: initialise the environment first");
    std::vector<State> cps;
    for(int k = 0; k < steps; ++k) {
        cps.push_back(state());
        grow_all();
    }
    const State final_state = state();
    Tape tape;
    struct Reset {
        Runtime* r;
        ~Reset() { r->tape = nullptr; }
    } reset{rt};
    rt->tape = &tape;
    const std::vector<DTensor> Adag0 = Adags;
    dA.assign(As.size(), DTensor());
    for(std::size_t i = 0; i < As.size(); ++i) dA[i] = DTensor::zeros(rt, As[i].st);
    State carried;  // adjoints of the environment after the move being reversed
    double E = 0.0;
    for(int k = steps - 1; k >= 0; --k) {
        set_state(cps[k]);
        tape.clear();
        tape.active = true;
        // iPEPS::updateAdags under the tape, so that the adjoint of A^dagger reaches A
        for(std::size_t i = 0; i < As.size(); ++i) Adags[i] = permute(adjoint(As[i]), 0, {3, 4, 2, 0, 1});
        grow_all();
        if(k == steps - 1) {
            compute_rdm();
            const std::vector<DTensor> terms = bond_terms(Jxy, Jz);
            tape.active = false;
            for(const DTensor& t : terms) {
                E += trace(t);
                DTensor dt = tape.adjoint_of(t);
                trace_bwd(dt, 1.0 / (Lx * Ly));
            }
            E /= (Lx * Ly);
        } else {
            tape.active = false;
            const State now = state();
            const std::vector<DTensor>* outs[8] = {&now.C1, &now.C2, &now.C3, &now.C4, &now.T1, &now.T2, &now.T3, &now.T4};
            const std::vector<DTensor>* adjs[8] = {&carried.C1, &carried.C2, &carried.C3, &carried.C4, &carried.T1, &carried.T2, &carried.T3, &carried.T4};
            for(int w = 0; w < 8; ++w)
                for(std::size_t i = 0; i < outs[w]->size(); ++i)
                    if((*adjs[w])[i].st) tape.accumulate((*outs[w])[i], (*adjs[w])[i]);
        }
        tape.backward();
        const State& in = cps[k];
        const std::vector<DTensor>* ins[8] = {&in.C1, &in.C2, &in.C3, &in.C4, &in.T1, &in.T2, &in.T3, &in.T4};
        std::vector<DTensor>* cad[8] = {&carried.C1, &carried.C2, &carried.C3, &carried.C4, &carried.T1, &carried.T2, &carried.T3, &carried.T4};
        for(int w = 0; w < 8; ++w) {
            cad[w]->assign(ins[w]->size(), DTensor());
            for(std::size_t i = 0; i < ins[w]->size(); ++i)
                if(const DTensor* d = tape.find((*ins[w])[i])) (*cad[w])[i] = *d;
        }
        for(std::size_t i = 0; i < As.size(); ++i)
            if(const DTensor* d = tape.find(As[i])) dA[i] = axpby(*d, dA[i], 1.0, 1.0);
    }
    tape.clear();
    Adags = Adag0;
    set_state(final_state);
    return E;
}

}  // namespace xa

