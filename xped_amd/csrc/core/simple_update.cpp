// Simple update on the GPU tensor layer (the reference functions it follows are listed in
// simple_update.hpp).  The contraction index lists restate the reference's network; twists are
// kept where the reference applies them (the identity for bosonic symmetries).
#include "simple_update.hpp"

#include <cmath>
#include <limits>
#include <string>

namespace xa {

namespace {
constexpr int L = 2;  // unit-cell extent in x and y

int idx(int x, int y)
{
    const int xm = ((x % L) + L) % L, ym = ((y % L) + L) % L;
    return xm * L + ym;
}

DTensor sqrt_w(const DTensor& w) { return diag_map(w, 0); }
DTensor inv_w(const DTensor& w) { return diag_map(w, 1); }

DTensor identity_on(const DTensor& w)
{
    // iPEPS::Id_weight_h / Id_weight_v (iPEPS.cpp:605-615): the identity on the weight's legs
    return identity(w.rt, w.st->unc_dom, w.st->unc_cod);
}

DTensor max_normalized(const DTensor& w)
{
    // weight * (1. / weight.maxNorm()) (TimePropagator.cpp:176, 210)
    const double m = max_norm(w);
    if(!(m > 0.0) || !std::isfinite(m)) throw Error(XH_EINVAL, "simple update: bond weight with max norm " + std::to_string(m));
    return scaled(w, 1.0 / m);
}
}  // namespace

WeightedPEPS init_weight_tensors(Runtime* rt, const std::vector<DTensor>& A, std::size_t D)
{
    if(A.size() != L * L) throw Error(XH_EINVAL, "simple update: expected 4 site tensors for the 2x2 unit cell");
    if(D == 0) throw Error(XH_EINVAL, "simple update: bond dimension D must be positive");
    WeightedPEPS p;
    p.D = D;
    p.Gs = A;
    for(const auto& a : A) {
        if(!a.st || a.st->rank != 2 || a.st->corank != 3) throw Error(XH_EINVAL, "simple update: site tensors must be Tensor<2,3>(left, top | right, bottom, phys)");
        const Qbasis& right = a.st->unc_cod[0];
        const Qbasis& top = a.st->unc_dom[1];
        p.whs.push_back(identity(rt, {right}, {right}));
        p.wvs.push_back(identity(rt, {top}, {top}));
    }
    return p;
}

DTensor apply_weights(const DTensor& A, const DTensor& wL, const DTensor& wT, const DTensor& wR, const DTensor& wB, int precision)
{
    DTensor t = contract(A, {1, -2, -3, -4, -5}, twist(wL, 1), {-1, 1}, 2, precision);
    t = contract(t, {-1, 1, -3, -4, -5}, twist(wT, 1), {-2, 1}, 2, precision);
    t = contract(t, {-1, -2, 1, -4, -5}, wR, {1, -3}, 2, precision);
    return contract(t, {-1, -2, -3, 1, -5}, wB, {1, -4}, 2, precision);
}

std::vector<DTensor> weighted_sites(const WeightedPEPS& p)
{
    std::vector<DTensor> A(L * L);
    for(int x = 0; x < L; ++x)
        for(int y = 0; y < L; ++y)
            A[idx(x, y)] = apply_weights(p.Gs[idx(x, y)], sqrt_w(p.whs[idx(x - 1, y)]), sqrt_w(p.wvs[idx(x, y)]), sqrt_w(p.whs[idx(x, y)]),
                                         sqrt_w(p.wvs[idx(x, y + 1)]), XH_FP64);
    return A;
}

TimePropagator::TimePropagator(Runtime* r, const std::vector<DTensor>& Hh, const std::vector<DTensor>& Hv, double dt_) : rt(r), dt(dt_)
{
    if(Hh.size() != L * L || Hv.size() != L * L) throw Error(XH_EINVAL, "simple update: expected 4 horizontal and 4 vertical bond Hamiltonians");
    for(const auto* H : {&Hh, &Hv})
        for(const auto& h : *H)
            if(!h.st || h.st->rank != 2 || h.st->corank != 2) throw Error(XH_EINVAL, "simple update: bond Hamiltonians must be Tensor<2,2>");
    // initU: U = H.mexp(-dt)
    for(const auto& h : Hh) Uh.push_back(mexp(h, -dt));
    for(const auto& h : Hv) Uv.push_back(mexp(h, -dt));
    spectrum_h.resize(L * L);
    spectrum_v.resize(L * L);
}

void TimePropagator::t_step(WeightedPEPS& psi)
{
    // the nearest-neighbour branch: horizontal then vertical bonds over the unique sites in
    // pattern order (coords(i) = (i % Lx, i / Lx)), then both again in reverse order
    const int n = L * L;
    for(int i = 0; i < n; ++i) t_step_h(psi, i % L, i / L);
    for(int i = 0; i < n; ++i) t_step_v(psi, i % L, i / L);
    for(int i = n - 1; i >= 0; --i) t_step_v(psi, i % L, i / L);
    for(int i = n - 1; i >= 0; --i) t_step_h(psi, i % L, i / L);
}

void TimePropagator::t_step_h(WeightedPEPS& psi, int x, int y)
{
    auto& G = psi.Gs;
    auto& wh = psi.whs;
    auto& wv = psi.wvs;
    const std::size_t all = std::numeric_limits<std::size_t>::max();
    DTensor left_full = apply_weights(G[idx(x, y)], wh[idx(x - 1, y)], wv[idx(x, y)], sqrt_w(wh[idx(x, y)]), wv[idx(x, y + 1)], precision);
    DTensor right_full = apply_weights(G[idx(x + 1, y)], sqrt_w(wh[idx(x, y)]), wv[idx(x + 1, y)], wh[idx(x + 1, y)], wv[idx(x + 1, y + 1)], precision);
    // (left, top, bottom | right, phys) = left * bond_l, (left, phys | top, right, bottom) = bond_r * right
    SVDResult l = tsvd(permute(left_full, -1, {0, 1, 3, 2, 4}), all, 0.0, false);
    DTensor bond_l = mul(l.S, l.Vdag, precision);
    SVDResult r = tsvd(permute(right_full, 0, {0, 4, 1, 2, 3}), all, 0.0, false);
    DTensor bond_r = mul(r.U, r.S, precision);
    DTensor bond = contract(bond_l, {-1, 1, -2}, bond_r, {1, -3, -4}, 2, precision);
    DTensor enlarged = contract(twist(twist(Uh[idx(x, y)], 0), 1), {1, 2, -1, -3}, bond, {-2, 1, 2, -4}, 2, precision);
    // SimpleUpdate::renormalize: bond.tSVD(D, 1e-14)
    SVDResult nb = tsvd(enlarged, psi.D, 1e-14, false);
    spectrum_h[idx(x, y)] = nb.S;
    wh[idx(x, y)] = max_normalized(nb.S);
    DTensor g = contract(l.U, {-1, -2, -4, 1}, nb.U, {-5, 1, -3}, 2, precision);
    G[idx(x, y)] = apply_weights(g, inv_w(wh[idx(x - 1, y)]), inv_w(wv[idx(x, y)]), identity_on(wh[idx(x, y)]), inv_w(wv[idx(x, y + 1)]), precision);
    g = contract(nb.Vdag, {-1, -5, 1}, r.Vdag, {1, -2, -3, -4}, 2, precision);
    G[idx(x + 1, y)] =
        apply_weights(g, identity_on(wh[idx(x, y)]), inv_w(wv[idx(x + 1, y)]), inv_w(wh[idx(x + 1, y)]), inv_w(wv[idx(x + 1, y + 1)]), precision);
    ++n_bond_updates;
}

void TimePropagator::t_step_v(WeightedPEPS& psi, int x, int y)
{
    auto& G = psi.Gs;
    auto& wh = psi.whs;
    auto& wv = psi.wvs;
    const std::size_t all = std::numeric_limits<std::size_t>::max();
    DTensor top_full = apply_weights(G[idx(x, y - 1)], wh[idx(x - 1, y - 1)], wv[idx(x, y - 1)], wh[idx(x, y - 1)], sqrt_w(wv[idx(x, y)]), precision);
    DTensor bottom_full = apply_weights(G[idx(x, y)], wh[idx(x - 1, y)], sqrt_w(wv[idx(x, y)]), wh[idx(x, y)], wv[idx(x, y + 1)], precision);
    // (left, top, right | bottom, phys) = top * bond_t, (top, phys | left, right, bottom) = bond_b * bottom
    SVDResult t = tsvd(permute(top_full, -1, {0, 1, 2, 3, 4}), all, 0.0, false);
    DTensor bond_t = mul(t.S, t.Vdag, precision);
    SVDResult b = tsvd(permute(bottom_full, 0, {1, 4, 0, 2, 3}), all, 0.0, false);
    DTensor bond_b = mul(b.U, b.S, precision);
    DTensor bond = contract(bond_t, {-1, 1, -2}, bond_b, {1, -3, -4}, 2, precision);
    DTensor enlarged = contract(twist(twist(Uv[idx(x, y - 1)], 0), 1), {1, 2, -1, -3}, bond, {-2, 1, 2, -4}, 2, precision);
    SVDResult nb = tsvd(enlarged, psi.D, 1e-14, false);
    spectrum_v[idx(x, y)] = nb.S;
    wv[idx(x, y)] = max_normalized(nb.S);
    DTensor g = contract(t.U, {-1, -2, -3, 1}, nb.U, {-5, 1, -4}, 2, precision);
    G[idx(x, y - 1)] =
        apply_weights(g, inv_w(wh[idx(x - 1, y - 1)]), inv_w(wv[idx(x, y - 1)]), inv_w(wh[idx(x, y - 1)]), identity_on(wv[idx(x, y)]), precision);
    g = contract(nb.Vdag, {-2, -5, 1}, b.Vdag, {1, -1, -3, -4}, 2, precision);
    G[idx(x, y)] = apply_weights(g, inv_w(wh[idx(x - 1, y)]), identity_on(wv[idx(x, y)]), inv_w(wh[idx(x, y)]), inv_w(wv[idx(x, y + 1)]), precision);
    ++n_bond_updates;
}

}  // namespace xa
