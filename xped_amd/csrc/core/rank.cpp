// Learned rank prediction for the randomized block SVD (see ctm.hpp).
#include <algorithm>
#include <cmath>

#include "ctm.hpp"

namespace xa {

namespace {
// the features of a sector that do not depend on the cut (f[0..16]) and the two logarithms of f[18]
struct FeatureBase {
    float f[RANK_NFEAT];
    double lg_glob, lg_last;
};
double lg10(double x) { return x > 0.0 ? std::log10(x) : -300.0; }
FeatureBase feature_base(const double* top, int ntop, int dim, double s_glob)
{
    // decades, scaled by 1/4 so that the inputs are O(1)
    FeatureBase b{};
    const double s1 = ntop > 0 ? top[0] : 0.0;
    b.f[0] = static_cast<float>(0.25 * std::max(-16.0, lg10(s1) - lg10(s_glob)));
    for(int j = 1; j < RANK_TOP; ++j) b.f[j] = static_cast<float>(0.25 * (j < ntop ? std::max(-16.0, lg10(top[j]) - lg10(s1)) : -16.0));
    b.f[16] = static_cast<float>(0.25 * std::log10(std::max(1, dim)));
    const double slast = ntop > 0 ? top[std::min(ntop, RANK_TOP) - 1] : 0.0;
    b.lg_glob = lg10(s_glob);
    b.lg_last = lg10(slast);
    return b;
}
void features_at(const FeatureBase& b, double log_cut, float* f)
{
    std::copy(b.f, b.f + 17, f);
    f[17] = static_cast<float>(0.25 * log_cut);
    f[18] = static_cast<float>(0.25 * std::max(-16.0, std::min(16.0, log_cut + b.lg_glob - b.lg_last)));
}
}  // namespace

void rank_features(const double* top, int ntop, int dim, double s_glob, double log_cut, float* f)
{
    features_at(feature_base(top, ntop, dim, s_glob), log_cut, f);
}

namespace {
void rchk(xh_ctx* ctx, int rc, const char* what)
{
    if(rc != XH_OK) throw Error(rc, std::string(what) + ": " + xh_last_error(ctx));
}
}  // namespace

RankPredictor::RankPredictor(xh_ctx* c, const float* params, int64_t nparams, int w, double m) : ctx(c), width(w), margin(m)
{
    rchk(ctx, xh_mlp_create(ctx, RANK_NFEAT, width, 1, &net), "rank predictor");
    int64_t n = 0;
    xh_mlp_num_params(net, &n);
    if(n != nparams) {
        xh_mlp_destroy(ctx, net);
        net = nullptr;
        throw Error(XH_EINVAL, "rank predictor: parameter count does not match the network shape");
    }
    rchk(ctx, xh_mlp_set_params(ctx, net, params), "rank predictor parameters");
}

RankPredictor::~RankPredictor()
{
    if(net) xh_mlp_destroy(ctx, net);
}

std::vector<double> RankPredictor::counts(const std::vector<std::vector<double>>& top, const std::vector<int>& dims, double s_glob, double log_cut) const
{
    return counts_grid(top, dims, s_glob, std::vector<double>{log_cut});
}

std::vector<double> RankPredictor::counts_grid(const std::vector<std::vector<double>>& top, const std::vector<int>& dims, double s_glob,
                                               const std::vector<double>& log_cuts) const
{
    // one batched forward for every (cut, sector) pair: rows cut-major
    const int nsec = static_cast<int>(top.size());
    const int64_t B = static_cast<int64_t>(nsec) * static_cast<int64_t>(log_cuts.size());
    std::vector<float> X(static_cast<std::size_t>(B) * RANK_NFEAT);
    // the logarithms of a sector's values once, then per cut its two cut features
    std::vector<FeatureBase> base(static_cast<std::size_t>(nsec));
    for(int q = 0; q < nsec; ++q) base[q] = feature_base(top[q].data(), static_cast<int>(top[q].size()), dims[q], s_glob);
    for(std::size_t c = 0; c < log_cuts.size(); ++c)
        for(int q = 0; q < nsec; ++q) features_at(base[q], log_cuts[c], X.data() + (c * nsec + q) * RANK_NFEAT);
    float *dX = nullptr, *dY = nullptr;
    rchk(ctx, xh_alloc(ctx, X.size() * sizeof(float), reinterpret_cast<void**>(&dX)), "alloc");
    rchk(ctx, xh_alloc(ctx, std::max<int64_t>(B, 1) * sizeof(float), reinterpret_cast<void**>(&dY)), "alloc");
    std::vector<float> y(static_cast<std::size_t>(B));
    int rc = xh_upload(ctx, dX, X.data(), X.size() * sizeof(float));
    if(rc == XH_OK) rc = xh_mlp_forward(ctx, net, dX, B, dY);
    if(rc == XH_OK) rc = xh_download(ctx, y.data(), dY, static_cast<std::size_t>(B) * sizeof(float));
    xh_free(ctx, dX);
    xh_free(ctx, dY);
    rchk(ctx, rc, "rank predictor forward");
    std::vector<double> out(static_cast<std::size_t>(B));
    for(int64_t i = 0; i < B; ++i) out[i] = std::min<double>(dims[i % nsec], std::max(0.0, std::expm1(static_cast<double>(y[i]))));
    return out;
}

namespace {
// the sketch the features are read from: the RANK_TOP leading singular values of every sector
RsvdOptions sketch_options()
{
    RsvdOptions o;
    o.oversample = 4;
    o.rank_factor = 1.0;
    o.power_iters = 2;
    return o;
}
std::map<Q, int> sketch_ranks(const TensorStructure& s)
{
    std::map<Q, int> sk;
    for(const Q& q : s.sector) sk[q] = RANK_TOP;
    return sk;
}
}  // namespace

std::map<Q, int> RankPredictor::predict(const DTensor& prod, std::size_t chi, std::uint64_t seed) const
{
    RandomizedSVD r = tsvd_randomized(prod, sketch_ranks(*prod.st), static_cast<std::size_t>(-1) / 2, 0.0, sketch_options(), seed);
    return ranks_from_spectra(*prod.st, r.res.spec, chi);
}

std::vector<std::map<Q, int>> RankPredictor::predict_batch(const std::vector<const DTensor*>& prods, std::size_t chi,
                                                           const std::vector<std::uint64_t>& seeds) const
{
    // the sketches of several tensors share grouped launches (tsvd_randomized_batch)
    std::vector<std::map<Q, int>> sk;
    for(const DTensor* p : prods) sk.push_back(sketch_ranks(*p->st));
    std::vector<RandomizedSVD> rs = tsvd_randomized_batch(prods, sk, static_cast<std::size_t>(-1) / 2, 0.0, sketch_options(), seeds);
    std::vector<std::map<Q, int>> out;
    for(std::size_t b = 0; b < prods.size(); ++b) out.push_back(ranks_from_spectra(*prods[b]->st, rs[b].res.spec, chi));
    return out;
}

std::map<Q, int> RankPredictor::ranks_from_spectra(const TensorStructure& s, const std::map<Q, std::vector<double>>& spec, std::size_t chi) const
{
    std::vector<std::vector<double>> top;
    std::vector<int> dims;
    std::vector<Q> qs;
    double s_glob = 0.0;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        const int dim = static_cast<int>(std::min(s.rows[i], s.cols[i]));
        if(dim == 0) continue;
        std::vector<double> t;
        auto it = spec.find(s.sector[i]);
        if(it != spec.end()) {
            t = it->second;
            std::sort(t.begin(), t.end(), std::greater<double>());
            if(t.size() > static_cast<std::size_t>(RANK_TOP)) t.resize(RANK_TOP);
        }
        if(!t.empty()) s_glob = std::max(s_glob, t[0]);
        top.push_back(std::move(t));
        dims.push_back(dim);
        qs.push_back(s.sector[i]);
    }
    std::map<Q, int> ranks;
    if(top.empty() || !(s_glob > 0.0)) return ranks;
    // the cut (in decades below the largest value) where the predicted total reaches chi: a coarse
    // grid over 14 decades, then a fine grid inside the crossing cell, each one batched forward
    const int nsec = static_cast<int>(qs.size());
    auto crossing = [&](double lo, double hi, int n, double& clo, double& chi_cut) {
        std::vector<double> grid(n + 1);
        for(int i = 0; i <= n; ++i) grid[i] = lo + (hi - lo) * i / n;
        const std::vector<double> c = counts_grid(top, dims, s_glob, grid);
        // totals fall as the cut rises; find the last grid point whose total still exceeds chi
        int last_above = -1;
        for(int i = 0; i <= n; ++i) {
            double tot = 0.0;
            for(int q = 0; q < nsec; ++q) tot += c[static_cast<std::size_t>(i) * nsec + q];
            if(tot > static_cast<double>(chi)) last_above = i;
        }
        clo = last_above < 0 ? lo : grid[last_above];
        chi_cut = last_above < 0 ? lo : (last_above >= n ? hi : grid[last_above + 1]);
    };
    double lo = -14.0, hi = 0.0;
    crossing(-14.0, 0.0, 280, lo, hi);
    crossing(lo, hi, 64, lo, hi);
    const std::vector<double> cnt = counts(top, dims, s_glob, hi);
    for(std::size_t q = 0; q < qs.size(); ++q) ranks[qs[q]] = static_cast<int>(std::ceil(margin * cnt[q]));
    return ranks;
}

}  // namespace xa
