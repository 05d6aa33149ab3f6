// Host-side integer structure engine (fusion trees, bases, tree-pair moves, plans).
#include "structure.hpp"

#include "../common/pool.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <numeric>
#include <sstream>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

namespace xa {

using xp::WorkerPool;
using xp::parallel_ranges;


// ---------------------------------------------------------------- FTree

FTree FTree::enlarge(const FTree& other) const
{
    FTree out;
    out.rank = rank + 1;
    for(int i = 0; i < rank; ++i) {
        out.dims[i] = dims[i];
        out.unc[i] = unc[i];
        out.dual[i] = dual[i];
    }
    out.dims[rank] = other.dims[0];
    out.unc[rank] = other.unc[0];
    out.dual[rank] = other.dual[0];
    out.compute_dim();
    for(int i = 0; i < n_inter(); ++i) out.inter[i] = inter[i];
    if(out.n_inter() > 0) out.inter[out.n_inter() - 1] = coupled;
    return out;
}

void FTree::compute_intermediates(const Sym& s)
{
    if(rank == 0) return;
    for(int i = 0; i < n_inter(); ++i)
        inter[i] = (i == 0) ? s.combine(unc[0], unc[1]).front() : s.combine(inter[i - 1], unc[i + 1]).front();
    if(rank < 2)
        coupled = unc[0];
    else if(rank == 2)
        coupled = s.combine(unc[0], unc[1]).front();
    else
        coupled = s.combine(inter[n_inter() - 1], unc[rank - 1]).front();
}

bool FTreeKeyLess::operator()(const FTree& a, const FTree& b) const
{
    if(a.rank != b.rank) return a.rank < b.rank;
    for(int i = 0; i < a.rank; ++i) {
        if(a.unc[i] != b.unc[i]) return a.unc[i] < b.unc[i];
    }
    for(int i = 0; i < a.n_inter(); ++i) {
        if(a.inter[i] != b.inter[i]) return a.inter[i] < b.inter[i];
    }
    if(a.coupled != b.coupled) return a.coupled < b.coupled;
    for(int i = 0; i < a.rank; ++i) {
        if(a.dims[i] != b.dims[i]) return a.dims[i] < b.dims[i];
        if(a.dual[i] != b.dual[i]) return a.dual[i] < b.dual[i];
    }
    return false;
}

// ---------------------------------------------------------------- Permutation

Permutation::Permutation(std::vector<int> p) : pi(std::move(p))
{
    pi_inv.assign(pi.size(), 0);
    std::vector<int> seen(pi.size(), 0);
    for(std::size_t i = 0; i < pi.size(); ++i) {
        if(pi[i] < 0 || pi[i] >= static_cast<int>(pi.size()) || seen[pi[i]]) throw std::invalid_argument("invalid permutation");
        seen[pi[i]] = 1;
        pi_inv[pi[i]] = static_cast<int>(i);
    }
}

bool Permutation::is_identity() const
{
    for(std::size_t i = 0; i < pi.size(); ++i)
        if(pi[i] != static_cast<int>(i)) return false;
    return true;
}

std::vector<int> Permutation::decompose() const
{
    // Bubble the target sequence into place with adjacent swaps; the swap positions are recorded
    // from the rightmost one down to the target slot, as Permutations.hpp:195-225 does.
    std::vector<int> a(pi.size());
    std::iota(a.begin(), a.end(), 0);
    std::vector<int> out;
    if(a == pi) return out;
    for(int i = 0; i < static_cast<int>(pi.size()); ++i) {
        int index = static_cast<int>(std::find(a.begin(), a.end(), pi[i]) - a.begin());
        if(index == 0) continue;
        for(int j = index - 1; j >= i; --j) {
            std::swap(a[j], a[j + 1]);
            out.push_back(j);
        }
    }
    return out;
}

// ---------------------------------------------------------------- Qbasis

void Qbasis::push_back(const Q& q, std::size_t inner)
{
    if(inner == 0) return;
    int i = find(q);
    if(i < 0) {
        QEntry e;
        e.q = q;
        e.num = curr_dim;
        e.dim = inner;
        FTree t;
        if(depth == 1) {
            t.rank = 1;
            t.unc[0] = q;
            t.coupled = q;
            t.dual[0] = conj_flag;
            t.dims[0] = static_cast<std::uint32_t>(inner);
            t.dim = inner;
        } else {
            t.rank = 0;
            t.dim = 1;
            t.coupled = q;
        }
        e.trees.push_back(t);
        data.push_back(e);
        sorted = false;
        tree_index.reset();
    } else {
        tree_index.reset();
        data[i].dim += inner;
        for(std::size_t j = i + 1; j < data.size(); ++j) data[j].num += inner;
        if(depth == 1) {
            data[i].trees[0].dim += inner;
            data[i].trees[0].dims[0] += static_cast<std::uint32_t>(inner);
        }
    }
    curr_dim += inner;
}

void Qbasis::push_back_trees(const Q& q, std::size_t inner, const std::vector<FTree>& t)
{
    if(find(q) >= 0) throw std::logic_error("Qbasis::push_back with trees: quantum number already present");
    QEntry e;
    e.q = q;
    e.num = curr_dim;
    e.dim = inner;
    e.trees = t;
    data.push_back(e);
    sorted = false;
    tree_index.reset();
    curr_dim += inner;
}

std::vector<std::size_t> Qbasis::sort()
{
    if(data.size() < 2) return {0};
    std::vector<std::size_t> idx(data.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](std::size_t a, std::size_t b) {
        if(conj_flag) return sym.compare(sym.conj(data[a].q), sym.conj(data[b].q));
        return sym.compare(data[a].q, data[b].q);
    });
    std::vector<QEntry> nd(data.size());
    for(std::size_t i = 0; i < data.size(); ++i) nd[i] = data[idx[i]];
    std::size_t acc = 0;
    for(auto& e : nd) {
        e.num = acc;
        acc += e.dim;
    }
    data = std::move(nd);
    sorted = true;
    tree_index.reset();
    return idx;
}

Qbasis Qbasis::conj() const
{
    Qbasis out(sym, 1);
    out.conj_flag = !conj_flag;
    for(const auto& e : data) {
        const FTree& old = e.trees.at(0);
        FTree t;
        t.rank = 1;
        t.unc[0] = sym.conj(e.q);
        t.coupled = sym.conj(e.q);
        t.dual[0] = !old.dual[0];
        t.dims[0] = old.dims[0];
        t.dim = old.dim;
        out.push_back_trees(sym.conj(e.q), e.dim, {t});
    }
    out.sort();
    return out;
}

Qbasis Qbasis::combine(const Qbasis& other) const
{
    Qbasis out(sym, depth + 1);
    // coupled q -> trees in discovery order; the final order is fixed by the sorts below
    std::vector<Q> qs;
    std::vector<std::vector<FTree>> trees;
    auto slot = [&](const Q& q) -> std::vector<FTree>& {
        for(std::size_t i = 0; i < qs.size(); ++i)
            if(qs[i] == q) return trees[i];
        qs.push_back(q);
        trees.emplace_back();
        return trees.back();
    };
    for(const auto& e1 : data) {
        for(const auto& e2 : other.data) {
            for(const Q& q : sym.combine(e1.q, e2.q)) {
                for(const auto& t1 : e1.trees)
                    for(const auto& t2 : e2.trees) {
                        // Qbasis.cpp:463-468 skips trees already present; distinct (q1, q2, t1, t2, q)
                        // always give distinct enlarged trees, so the linear search is omitted.
                        FTree tt = t1.enlarge(t2);
                        tt.coupled = q;
                        slot(q).push_back(tt);
                    }
            }
        }
    }
    for(auto& v : trees) std::sort(v.begin(), v.end(), [](const FTree& a, const FTree& b) { return a.less(b); });
    for(std::size_t i = 0; i < qs.size(); ++i) {
        std::size_t inner = 0;
        for(const auto& t : trees[i]) inner += t.dim;
        out.push_back_trees(qs[i], inner, trees[i]);
    }
    out.sort();
    return out;
}

Qbasis Qbasis::forget_history() const
{
    Qbasis out(sym, 1);
    for(const auto& e : data) out.push_back(e.q, e.dim);
    return out;
}

std::uint64_t FTree::hash() const
{
    // hashes exactly the fields that same() compares
    auto mix = [](std::uint64_t h, std::uint64_t v) {
        h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        return h;
    };
    auto qv = [](const Q& q) { return (static_cast<std::uint64_t>(static_cast<std::uint32_t>(q.v[0])) << 32) | static_cast<std::uint32_t>(q.v[1]); };
    std::uint64_t h = mix(static_cast<std::uint64_t>(rank), qv(coupled));
    for(int i = 0; i < rank; ++i) h = mix(mix(h, qv(unc[i])), dims[i]);
    for(int i = 0; i < n_inter(); ++i) h = mix(h, qv(inter[i]));
    return h;
}

std::size_t Qbasis::left_offset(const FTree& t) const
{
    // Qbasis::leftOffset (Qbasis.cpp:208-225): sum of the dims of the trees before t in its
    // sector, found through a hash index instead of a linear scan
    int i = find(t.coupled);
    if(i < 0) throw std::runtime_error("Qbasis::leftOffset: coupled sector not present");
    if(!tree_index) {
        auto idx = std::make_shared<TreeIndex>(data.size());
        for(std::size_t s = 0; s < data.size(); ++s) {
            std::size_t off = 0;
            (*idx)[s].reserve(data[s].trees.size());
            for(std::size_t k = 0; k < data[s].trees.size(); ++k) {
                (*idx)[s].emplace(data[s].trees[k].hash(), std::make_pair(k, off));
                off += data[s].trees[k].dim;
            }
        }
        tree_index = idx;
    }
    const auto& m = (*tree_index)[i];
    auto range = m.equal_range(t.hash());
    for(auto it = range.first; it != range.second; ++it)
        if(data[i].trees[it->second.first].same(t)) return it->second.second;
    throw std::runtime_error("Qbasis::leftOffset: tree not present");
}

std::pair<std::size_t, std::size_t> Qbasis::tree_pos(const FTree& t) const
{
    (void)left_offset(t);  // builds the index and checks presence
    const int i = find(t.coupled);
    const auto& m = (*tree_index)[i];
    auto range = m.equal_range(t.hash());
    for(auto it = range.first; it != range.second; ++it)
        if(data[i].trees[it->second.first].same(t)) return it->second;
    throw std::runtime_error("Qbasis::tree_pos: tree not present");
}

bool Qbasis::equals(const Qbasis& o) const
{
    if(data.size() != o.data.size()) return false;
    for(std::size_t i = 0; i < data.size(); ++i) {
        const auto& a = data[i];
        const auto& b = o.data[i];
        if(a.q != b.q || a.num != b.num || a.dim != b.dim || a.trees.size() != b.trees.size()) return false;
        for(std::size_t t = 0; t < a.trees.size(); ++t)
            if(!a.trees[t].same(b.trees[t])) return false;
    }
    return true;
}

namespace {

// The coupled basis of a list of legs depends on their dimensions only through the tree and
// entry dimensions: the trees, their order and the entry order follow from the charges (combine
// sorts trees by charges and entries by q).  Coupled bases are therefore kept per charge pattern
// and re-dimensioned for legs with the same charges and other dimensions.
std::string charge_pattern(const Sym& s, const std::vector<Qbasis>& legs)
{
    std::string k;
    auto put = [&k](long long v) { k.append(reinterpret_cast<const char*>(&v), sizeof(v)); };
    put(s.kind[0]);
    put(s.kind[1]);
    put(static_cast<long long>(legs.size()));
    for(const auto& b : legs) {
        put(b.conj_flag ? 1 : 0);
        put(b.depth);
        put(static_cast<long long>(b.data.size()));
        for(const auto& e : b.data) {
            put(e.q.v[0]);
            put(e.q.v[1]);
            put(e.trees.empty() ? 0 : (e.trees[0].dual[0] ? 1 : 0));
        }
    }
    return k;
}

std::mutex pattern_mutex;
std::unordered_map<std::string, std::shared_ptr<const Qbasis>> pattern_cache;

Qbasis redimension(const Qbasis& tmpl, const std::vector<Qbasis>& legs)
{
    Qbasis b = tmpl;
    std::size_t acc = 0;
    for(auto& e : b.data) {
        std::size_t d = 0;
        for(auto& t : e.trees) {
            for(int i = 0; i < t.rank; ++i) t.dims[i] = static_cast<std::uint32_t>(legs[static_cast<std::size_t>(i)].inner_dim(t.unc[i]));
            t.compute_dim();
            d += t.dim;
        }
        e.dim = d;
        e.num = acc;
        acc += d;
    }
    b.curr_dim = acc;
    b.tree_index.reset();
    return b;
}

}  // namespace

Qbasis build_fusion_tree(const Sym& s, const std::vector<Qbasis>& uncoupled)
{
    if(uncoupled.empty()) {
        Qbasis b(s, 0);
        b.push_back(s.vacuum(), 1);
        return b;
    }
    if(uncoupled.size() == 1) return uncoupled[0];
    bool depth1 = true;
    for(const auto& b : uncoupled) depth1 = depth1 && b.depth == 1;
    std::string key;
    if(depth1) {
        key = charge_pattern(s, uncoupled);
        std::lock_guard<std::mutex> lk(pattern_mutex);
        auto it = pattern_cache.find(key);
        if(it != pattern_cache.end()) return redimension(*it->second, uncoupled);
    }
    Qbasis b = uncoupled[0];
    for(std::size_t i = 1; i < uncoupled.size(); ++i) b = b.combine(uncoupled[i]);
    if(depth1) {
        std::lock_guard<std::mutex> lk(pattern_mutex);
        if(pattern_cache.size() > 8192) pattern_cache.clear();
        auto tmpl = std::make_shared<Qbasis>(b);
        tmpl->tree_index.reset();
        pattern_cache.emplace(key, std::move(tmpl));
    }
    return b;
}

std::shared_ptr<const Qbasis> coupled_basis(const Sym& s, const std::vector<Qbasis>& uncoupled)
{
    bool depth1 = uncoupled.size() >= 2;
    for(const auto& b : uncoupled) depth1 = depth1 && b.depth == 1;
    if(!depth1) return std::make_shared<const Qbasis>(build_fusion_tree(s, uncoupled));
    std::string key = charge_pattern(s, uncoupled);
    for(const auto& b : uncoupled)
        for(const auto& e : b.data) {
            const long long d = static_cast<long long>(e.dim);
            key.append(reinterpret_cast<const char*>(&d), sizeof(d));
        }
    // bounded by the trees held: a full cache is dropped (its bases stay alive in their structures)
    static std::mutex basis_mutex;
    static std::unordered_map<std::string, std::shared_ptr<const Qbasis>> basis_cache;
    static std::size_t held_trees = 0;
    constexpr std::size_t max_held_trees = std::size_t(1) << 21;
    {
        std::lock_guard<std::mutex> lk(basis_mutex);
        auto it = basis_cache.find(key);
        if(it != basis_cache.end()) return it->second;
    }
    auto b = std::make_shared<const Qbasis>(build_fusion_tree(s, uncoupled));
    std::size_t nt = 0;
    for(const auto& e : b->data) nt += e.trees.size();
    std::lock_guard<std::mutex> lk(basis_mutex);
    if(held_trees + nt > max_held_trees) {
        basis_cache.clear();
        held_trees = 0;
    }
    if(basis_cache.emplace(std::move(key), b).second) held_trees += nt;
    return b;
}

// ---------------------------------------------------------------- tree moves

TreeMap tree_swap(const Sym& s, const FTree& t, int pos)
{
    TreeMap out;
    const int R = t.rank;
    FTree tree = t;
    std::swap(tree.unc[pos], tree.unc[pos + 1]);
    std::swap(tree.dual[pos], tree.dual[pos + 1]);
    std::swap(tree.dims[pos], tree.dims[pos + 1]);
    if(pos == 0) {
        const Q& ql = t.unc[0];
        const Q& qr = t.unc[1];
        const Q& qf = (R == 2) ? t.coupled : t.inter[0];
        double c = s.coeff_swap(ql, qr, qf);
        if(std::abs(c) < TREE_EPS) return out;
        out.emplace(tree, c);
        return out;
    }
    const Q q1 = (pos == 1) ? t.unc[0] : t.inter[pos - 2];
    const Q q2 = t.unc[pos];
    const Q q3 = t.unc[pos + 1];
    const Q Qt = (pos == R - 2) ? t.coupled : t.inter[pos];
    const Q Q12 = t.inter[pos - 1];
    for(const Q& Q31 : s.combine(q3, q1)) {
        tree.inter[pos - 1] = Q31;
        double cgc = s.coeff_swap(Q12, q3, Qt) * s.coeff_recouple(q3, q1, q2, Qt, Q31, Q12) * s.coeff_swap(q3, q1, Q31);
        if(std::abs(cgc) < TREE_EPS) continue;
        out.emplace(tree, cgc);
    }
    return out;
}

TreeMap tree_permute(const Sym& s, const FTree& t, const Permutation& p)
{
    TreeMap out;
    out.emplace(t, 1.0);
    for(int pos : p.decompose()) {
        TreeMap tmp;
        for(const auto& [tree, c] : out) {
            for(const auto& [tree2, c2] : tree_swap(s, tree, pos)) {
                auto it = tmp.find(tree2);
                if(it == tmp.end()) {
                    tmp.emplace(tree2, c * c2);
                } else if(std::abs(it->second + c * c2) < TREE_EPS) {
                    tmp.erase(it);
                } else {
                    it->second += c * c2;
                }
            }
        }
        out.swap(tmp);
    }
    return out;
}

PairMap pair_turn_right(const Sym& s, const FTree& t1, const FTree& t2)
{
    const int R = t1.rank, C = t2.rank;
    if(C < 1) throw std::logic_error("turn_right needs a codomain leg");
    FTree t1p;
    t1p.rank = R + 1;
    for(int i = 0; i < R; ++i) {
        t1p.unc[i] = t1.unc[i];
        t1p.dual[i] = t1.dual[i];
        t1p.dims[i] = t1.dims[i];
    }
    t1p.unc[R] = s.conj(t2.unc[C - 1]);
    for(int i = 0; i < t1.n_inter(); ++i) t1p.inter[i] = t1.inter[i];
    if(R + 1 > 2) t1p.inter[t1p.n_inter() - 1] = t1.coupled;
    t1p.dual[R] = !t2.dual[C - 1];
    Q a;
    if(C == 1)
        a = s.vacuum();
    else if(C == 2)
        a = t2.unc[0];
    else
        a = t2.inter[t2.n_inter() - 1];
    t1p.coupled = a;
    t1p.dims[R] = t2.dims[C - 1];
    t1p.compute_dim();
    FTree t2p;
    t2p.rank = C - 1;
    for(int i = 0; i < C - 1; ++i) {
        t2p.unc[i] = t2.unc[i];
        t2p.dual[i] = t2.dual[i];
        t2p.dims[i] = t2.dims[i];
    }
    for(int i = 0; i < t2p.n_inter(); ++i) t2p.inter[i] = t2.inter[i];
    t2p.coupled = a;
    t2p.compute_dim();
    const Q b = t2.unc[C - 1];
    const Q c = t2.coupled;
    double coeff = s.coeff_turn(a, b, c);
    if(t2.dual[C - 1]) coeff *= s.coeff_FS(s.conj(b));
    PairMap out;
    if(std::abs(coeff) < TREE_EPS) return out;
    out.emplace(TreePair{t1p, t2p}, coeff);
    return out;
}

PairMap pair_turn_left(const Sym& s, const FTree& t1, const FTree& t2)
{
    PairMap out;
    for(const auto& [pr, c] : pair_turn_right(s, t2, t1)) out.emplace(TreePair{pr.second, pr.first}, c);
    return out;
}

PairMap pair_turn(const Sym& s, int shift, const FTree& t1, const FTree& t2)
{
    PairMap out;
    if(shift == 0) {
        out.emplace(TreePair{t1, t2}, 1.0);
        return out;
    }
    if(shift > 0) {
        for(const auto& [pr, c1] : pair_turn_left(s, t1, t2))
            for(const auto& [pr2, c2] : pair_turn(s, shift - 1, pr.first, pr.second)) out.emplace(pr2, c1 * c2);
        return out;
    }
    for(const auto& [pr, c1] : pair_turn_right(s, t1, t2))
        for(const auto& [pr2, c2] : pair_turn(s, shift + 1, pr.first, pr.second)) out.emplace(pr2, c1 * c2);
    return out;
}

// abelian bosonic fast path (treepair.cpp:147-168): one output pair, coefficient 1
static TreePair abelian_pair_image(const Sym& s, int shift, const FTree& t1, const FTree& t2, const Permutation& p)
{
    const int R = t1.rank, C = t2.rank;
    const int newR = R - shift, newC = C + shift;
    FTree od, oc;
    od.rank = newR;
    oc.rank = newC;
    for(int r = 0; r < newR; ++r) {
        int src = p.pi[r];
        od.unc[r] = src < R ? t1.unc[src] : s.conj(t2.unc[src - R]);
        od.dims[r] = src < R ? t1.dims[src] : t2.dims[src - R];
        od.dual[r] = src < R ? t1.dual[src] : !t2.dual[src - R];
    }
    for(int c = 0; c < newC; ++c) {
        int src = p.pi[c + newR];
        oc.unc[c] = src < R ? s.conj(t1.unc[src]) : t2.unc[src - R];
        oc.dims[c] = src < R ? t1.dims[src] : t2.dims[src - R];
        oc.dual[c] = src < R ? !t1.dual[src] : t2.dual[src - R];
    }
    od.compute_dim();
    od.compute_intermediates(s);
    oc.compute_dim();
    oc.compute_intermediates(s);
    return TreePair{od, oc};
}

PairMap pair_permute(const Sym& s, int shift, const FTree& t1, const FTree& t2, const Permutation& p)
{
    const int R = t1.rank, C = t2.rank;
    const int newR = R - shift, newC = C + shift;
    if(!s.any_non_abelian() && !s.any_fermionic()) {
        PairMap out;
        out.emplace(abelian_pair_image(s, shift, t1, t2, p), 1.0);
        return out;
    }
    // turn<-CoRank>() reverses the order of the codomain legs, so the permutation is corrected first
    const int N = R + C;
    std::vector<int> pi_id(N);
    std::iota(pi_id.begin(), pi_id.end(), 0);
    for(int i = 0; i < C; ++i) pi_id[i + R] = (C - 1) - i + R;
    std::vector<int> pi_tmp(N);
    for(int i = 0; i < N; ++i) pi_tmp[i] = pi_id[p.pi[i]];
    std::vector<int> pi_corr(N);
    for(int i = 0; i < newR; ++i) pi_corr[i] = pi_tmp[i];
    for(int i = 0; i < newC; ++i) pi_corr[i + newR] = pi_tmp[(newC - 1) - i + newR];
    Permutation pc(pi_corr);
    PairMap out;
    for(const auto& [pr, c1] : pair_turn(s, -C, t1, t2)) {
        for(const auto& [t1pp, c2] : tree_permute(s, pr.first, pc)) {
            for(const auto& [fin, c3] : pair_turn(s, newC, t1pp, pr.second)) {
                auto it = out.find(fin);
                if(it == out.end())
                    out.emplace(fin, c1 * c2 * c3);
                else
                    it->second += c1 * c2 * c3;
            }
        }
    }
    return out;
}

// ---------------------------------------------------------------- tensor structure

static void finish_layout(TensorStructure& t)
{
    std::size_t curr = 0;
    t.offsets.clear();
    t.rows.clear();
    t.cols.clear();
    for(const Q& q : t.sector) {
        t.offsets.push_back(curr);
        std::size_t r = t.dom().inner_dim(q), c = t.cod().inner_dim(q);
        t.rows.push_back(r);
        t.cols.push_back(c);
        curr += r * c;
    }
    t.size = curr;
}

StructPtr TensorStructure::create(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c)
{
    auto t = std::make_shared<TensorStructure>();
    t->sym = s;
    t->rank = static_cast<int>(d.size());
    t->corank = static_cast<int>(c.size());
    if(t->rank > MAXR || t->corank > MAXR) throw std::invalid_argument("too many legs");
    t->unc_dom = d;
    t->unc_cod = c;
    t->dom_p = coupled_basis(s, d);
    t->cod_p = coupled_basis(s, c);
    for(const auto& e : t->dom().data)
        if(t->cod().present(e.q)) t->sector.push_back(e.q);
    finish_layout(*t);
    return t;
}

StructPtr TensorStructure::create_pushed(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c, const std::vector<Q>& sectors)
{
    auto t = std::make_shared<TensorStructure>();
    t->sym = s;
    t->rank = static_cast<int>(d.size());
    t->corank = static_cast<int>(c.size());
    t->unc_dom = d;
    t->unc_cod = c;
    t->dom_p = coupled_basis(s, d);
    t->cod_p = coupled_basis(s, c);
    t->sector = sectors;
    finish_layout(*t);
    return t;
}

static void permuted_bases(const TensorStructure& in, int shift, const std::vector<int>& p, bool split, std::vector<Qbasis>& nd,
                           std::vector<Qbasis>& nc)
{
    const int R = in.rank;
    const int newR = in.rank - shift, newC = in.corank + shift;
    nd.clear();
    nc.clear();
    if(split) {
        for(int i = 0; i < newR; ++i) nd.push_back(in.unc_dom[p[i]]);
        for(int i = 0; i < newC; ++i) nc.push_back(in.unc_cod[p[i + newR] - R]);
        return;
    }
    for(int i = 0; i < newR; ++i) nd.push_back(p[i] >= R ? in.unc_cod[p[i] - R].conj() : in.unc_dom[p[i]]);
    for(int i = 0; i < newC; ++i) nc.push_back(p[i + newR] >= R ? in.unc_cod[p[i + newR] - R] : in.unc_dom[p[i + newR]].conj());
}

static void check_perm(const TensorStructure& in, int shift, const std::vector<int>& p)
{
    const int N = in.rank + in.corank;
    if(static_cast<int>(p.size()) != N) throw std::invalid_argument("permutation length does not match the tensor rank");
    if(shift > in.rank || -shift > in.corank) throw std::invalid_argument("invalid shift for permute");
    Permutation chk(p);  // validates
    (void)chk;
}

static bool is_split_case(const TensorStructure& in, int shift, const std::vector<int>& p)
{
    // Tensor.cpp:335-355: shift 0 and no leg crosses between domain and codomain
    if(shift != 0) return false;
    for(int i = 0; i < in.rank; ++i)
        if(p[i] >= in.rank) return false;
    for(int i = in.rank; i < in.rank + in.corank; ++i)
        if(p[i] < in.rank) return false;
    return true;
}

StructPtr permuted_structure(const TensorStructure& in, int shift, const std::vector<int>& p)
{
    check_perm(in, shift, p);
    std::vector<Qbasis> nd, nc;
    permuted_bases(in, shift, p, is_split_case(in, shift, p), nd, nc);
    return TensorStructure::create(in.sym, nd, nc);
}

SymbolicPermute make_symbolic_permute(const TensorStructure& in, const TensorStructure& out, int shift, const std::vector<int>& p)
{
    check_perm(in, shift, p);
    const int R = in.rank, C = in.corank, N = R + C;
    const bool split = is_split_case(in, shift, p);
    const bool abelian = !in.sym.any_non_abelian() && !in.sym.any_fermionic();
    SymbolicPermute sp;
    sp.nlegs = N;
    sp.new_rank = R - shift;
    // every output subblock is one task, so each output element is written exactly once
    int64_t ntasks = 0;
    for(std::size_t s = 0; s < out.sector.size(); ++s) {
        const Q& q = out.sector[s];
        sp.task_base.push_back(static_cast<std::int32_t>(ntasks));
        sp.ncod.push_back(static_cast<std::int32_t>(out.cod().tree(q).size()));
        ntasks += static_cast<int64_t>(out.dom().tree(q).size()) * out.cod().tree(q).size();
    }
    if(ntasks > 0x7fffffff) throw std::length_error("permute: too many subblocks");

    std::vector<int> pdv(p.begin(), p.begin() + R), pcv;
    for(int i = R; i < N; ++i) pcv.push_back(p[i] - R);
    Permutation P(p);
    std::unique_ptr<Permutation> Pd, Pc;
    if(split) {
        Pd = std::make_unique<Permutation>(pdv);
        Pc = std::make_unique<Permutation>(pcv);
    }
    // the source subblocks, in the reference's order (input sector, then row tree, then column
    // tree), are cut into contiguous ranges that host threads map to their tasks; the ranges are
    // merged in order, so every task keeps its sources in that order
    struct Work {
        std::int32_t sector, row_tree;
    };
    std::vector<Work> work;
    for(std::size_t i = 0; i < in.sector.size(); ++i) {
        const auto& dtrees = in.dom().tree(in.sector[i]);
        for(std::size_t a = 0; a < dtrees.size(); ++a) work.push_back(Work{static_cast<std::int32_t>(i), static_cast<std::int32_t>(a)});
    }
    // build the lazily created tree indices before threads share the bases
    for(const Qbasis* b : {&out.dom(), &out.cod()})
        if(!b->data.empty() && !b->data[0].trees.empty()) (void)b->left_offset(b->data[0].trees[0]);
    using Contrib = std::vector<std::pair<std::int32_t, SymbolicPermute::Src>>;
    auto run_range = [&](std::size_t w0, std::size_t w1, Contrib& res) {
        for(std::size_t w = w0; w < w1; ++w) {
            const std::int32_t i = work[w].sector;
            const Q& q = in.sector[i];
            const FTree& t1 = in.dom().tree(q)[work[w].row_tree];
            const auto& ctrees = in.cod().tree(q);
            for(std::size_t b2 = 0; b2 < ctrees.size(); ++b2) {
                const FTree& t2 = ctrees[b2];
                auto add = [&](const FTree& a, const FTree& b, double c) {
                    if(std::abs(c) < PERMUTE_SKIP) return;
                    const int so = out.find_sector(a.coupled);
                    if(so < 0) throw std::logic_error("permute: output sector missing");
                    const std::size_t ia = out.dom().tree_pos(a).first, ib = out.cod().tree_pos(b).first;
                    const std::int32_t task = sp.task_base[so] + static_cast<std::int32_t>(ia) * sp.ncod[so] + static_cast<std::int32_t>(ib);
                    res.emplace_back(task, SymbolicPermute::Src{i, work[w].row_tree, static_cast<std::int32_t>(b2), c});
                };
                if(abelian) {
                    // one image with coefficient 1, whether or not the permutation splits
                    const TreePair im = abelian_pair_image(in.sym, shift, t1, t2, P);
                    add(im.first, im.second, 1.0);
                } else if(split) {
                    TreeMap pd = tree_permute(in.sym, t1, *Pd);
                    TreeMap pc = tree_permute(in.sym, t2, *Pc);
                    for(const auto& [a, ca] : pd)
                        for(const auto& [b, cb] : pc) add(a, b, ca * cb);
                } else {
                    for(const auto& [pr, c] : pair_permute(in.sym, shift, t1, t2, P)) add(pr.first, pr.second, c);
                }
            }
        }
    };
    const std::size_t nw = work.size();
    const std::size_t nthreads = ntasks < 4096 ? 1 : std::min<std::size_t>(WorkerPool::get().size(), std::max<std::size_t>(1, nw / 4));
    std::vector<Contrib> parts(nthreads);
    if(nthreads == 1)
        run_range(0, nw, parts[0]);
    else
        WorkerPool::get().run(nthreads, [&](std::size_t k) { run_range(nw * k / nthreads, nw * (k + 1) / nthreads, parts[k]); });
    sp.src_begin.assign(static_cast<std::size_t>(ntasks), 0);
    sp.nsrc.assign(static_cast<std::size_t>(ntasks), 0);
    for(const auto& part : parts)
        for(const auto& [t, src] : part) ++sp.nsrc[static_cast<std::size_t>(t)];
    std::size_t total = 0;
    for(std::size_t t = 0; t < sp.nsrc.size(); ++t) {
        sp.src_begin[t] = static_cast<std::int32_t>(total);
        total += static_cast<std::size_t>(sp.nsrc[t]);
    }
    if(total > 0x7fffffff) throw std::length_error("permute: too many sources");
    sp.sources.resize(total);
    std::vector<std::int32_t> fill(sp.nsrc.size(), 0);
    for(const auto& part : parts)
        for(const auto& [t, src] : part) sp.sources[static_cast<std::size_t>(sp.src_begin[t] + fill[t]++)] = src;
    return sp;
}

namespace {
// left offsets of the trees of sector q in a coupled basis (Qbasis::leftOffset for every tree)
std::vector<std::size_t> tree_offsets(const Qbasis& b, const Q& q)
{
    const auto& trees = b.tree(q);
    std::vector<std::size_t> off(trees.size());
    std::size_t o = 0;
    for(std::size_t k = 0; k < trees.size(); ++k) {
        off[k] = o;
        o += trees[k].dim;
    }
    return off;
}
}  // namespace


namespace {
thread_local std::vector<PermuteTask> spare_tasks;
thread_local std::vector<PermuteSource> spare_sources;
}  // namespace

void recycle_plan_tables(PermutePlan& plan)
{
    if(plan.tasks.capacity() > spare_tasks.capacity()) spare_tasks = std::move(plan.tasks);
    if(plan.sources.capacity() > spare_sources.capacity()) spare_sources = std::move(plan.sources);
    plan.tasks = std::vector<PermuteTask>();
    plan.sources = std::vector<PermuteSource>();
}

PermutePlan numeric_permute_plan(const SymbolicPermute& sp, const TensorStructure& in, const StructPtr& outp, const std::vector<int>& p)
{
    const TensorStructure& out = *outp;
    const int R = in.rank, N = sp.nlegs, newR = sp.new_rank;
    if(N != in.rank + in.corank || sp.task_base.size() != out.sector.size()) throw std::logic_error("permute: symbolic plan does not match the structure");
    const Permutation Pfull(p);
    PermutePlan plan;
    plan.nlegs = N;
    plan.new_rank = newR;
    plan.out = outp;
    // every field of every task and source is written below, into recycled memory when available
    plan.tasks = std::move(spare_tasks);
    plan.sources = std::move(spare_sources);
    spare_tasks = std::vector<PermuteTask>();
    spare_sources = std::vector<PermuteSource>();
    plan.tasks.resize(sp.src_begin.size());
    for(std::size_t s = 0; s < out.sector.size(); ++s)
        if(static_cast<std::int32_t>(out.cod().tree(out.sector[s]).size()) != sp.ncod[s]) throw std::logic_error("permute: symbolic plan does not match the structure");
    // source stride of output leg k = product of the output extents of the source legs before p[k]
    std::array<std::array<int, 2 * MAXR>, 2 * MAXR> before{};
    std::array<int, 2 * MAXR> nbefore{};
    for(int k = 0; k < N; ++k) {
        const int m = p[k];
        const int lo = m < R ? 0 : R;
        for(int mm = lo; mm < m; ++mm) before[k][nbefore[k]++] = Pfull.pi_inv[mm];
    }
    // one work row per (output sector, domain tree): rows balance the host workers far better
    // than sectors, whose task counts differ by orders of magnitude
    struct Row {
        std::uint32_t s, a;
        std::size_t roff;
    };
    std::vector<Row> rows;
    for(std::size_t s = 0; s < out.sector.size(); ++s) {
        const auto& dtrees = out.dom().tree(out.sector[s]);
        std::size_t roff = 0;
        for(std::size_t a = 0; a < dtrees.size(); ++a) {
            rows.push_back(Row{static_cast<std::uint32_t>(s), static_cast<std::uint32_t>(a), roff});
            roff += dtrees[a].dim;
        }
    }
    std::vector<std::size_t> moved(rows.size(), 0);
    const std::size_t min_rows = plan.tasks.size() >= 16384 ? std::max<std::size_t>(1, rows.size() * 4096 / plan.tasks.size()) : rows.size() + 1;
    parallel_ranges(rows.size(), min_rows, [&](std::size_t r0, std::size_t r1) {
        for(std::size_t r = r0; r < r1; ++r) {
            const std::size_t s = rows[r].s, a = rows[r].a, roff = rows[r].roff;
            const Q& q = out.sector[s];
            const FTree& t1 = out.dom().tree(q)[a];
            const auto& ctrees = out.cod().tree(q);
            std::size_t coff = 0;
            for(std::size_t b = 0; b < ctrees.size(); ++b) {
                const FTree& t2 = ctrees[b];
                const std::size_t id = static_cast<std::size_t>(sp.task_base[s]) + a * ctrees.size() + b;
                PermuteTask& task = plan.tasks[id];
                task.dst_off = static_cast<std::int64_t>(out.offsets[s] + roff + coff * out.rows[s]);
                task.dst_ld = static_cast<std::int64_t>(out.rows[s]);
                task.rows = static_cast<std::int64_t>(t1.dim);
                task.cols = static_cast<std::int64_t>(t2.dim);
                for(int k = 0; k < N; ++k) task.ext[k] = k < newR ? t1.dims[k] : t2.dims[k - newR];
                for(int k = 0; k < N; ++k) {
                    std::int64_t st = 1;
                    for(int u = 0; u < nbefore[k]; ++u) st *= task.ext[before[k][u]];
                    task.sstride[k] = st;
                    task.is_col[k] = p[k] < R ? 0 : 1;
                }
                task.src_begin = sp.src_begin[id];
                task.nsrc = sp.nsrc[id];
                moved[r] += t1.dim * t2.dim;
                coff += t2.dim;
            }
        }
    });
    for(std::size_t m : moved) plan.moved_elems += m;
    std::vector<std::vector<std::size_t>> in_roff(in.sector.size()), in_coff(in.sector.size());
    for(std::size_t i = 0; i < in.sector.size(); ++i) {
        in_roff[i] = tree_offsets(in.dom(), in.sector[i]);
        in_coff[i] = tree_offsets(in.cod(), in.sector[i]);
    }
    plan.sources.resize(sp.sources.size());
    parallel_ranges(sp.sources.size(), 16384, [&](std::size_t k0, std::size_t k1) {
        for(std::size_t k = k0; k < k1; ++k) {
            const auto& src = sp.sources[k];
            const std::size_t i = static_cast<std::size_t>(src.sector);
            if(i >= in.sector.size() || static_cast<std::size_t>(src.t1) >= in_roff[i].size() || static_cast<std::size_t>(src.t2) >= in_coff[i].size())
                throw std::logic_error("permute: symbolic plan does not match the structure");
            plan.sources[k] = PermuteSource{static_cast<std::int64_t>(in.offsets[i] + in_roff[i][src.t1] + in_coff[i][src.t2] * in.rows[i]),
                                            static_cast<std::int64_t>(in.rows[i]), src.coeff};
        }
    });
    return plan;
}

PermutePlan make_permute_plan(const TensorStructure& in, int shift, const std::vector<int>& p)
{
    StructPtr out = permuted_structure(in, shift, p);
    return numeric_permute_plan(make_symbolic_permute(in, *out, shift, p), in, out, p);
}

StructPtr permute_adj_structure(const TensorStructure& y, int shift, const std::vector<int>& p)
{
    const int N = y.rank + y.corank;
    if(static_cast<int>(p.size()) != N) throw std::invalid_argument("permute_adj: permutation length does not match the tensor rank");
    const int R = y.rank + shift;  // rank of X
    if(R < 0 || R > N) throw std::invalid_argument("permute_adj: invalid shift");
    const Permutation P(p);
    std::vector<Qbasis> nd, nc;
    for(int m = 0; m < N; ++m) {
        const int k = P.pi_inv[m];  // Y leg that reads X leg m
        const bool y_cod = k >= y.rank;
        const Qbasis& leg = y_cod ? y.unc_cod[k - y.rank] : y.unc_dom[k];
        if(m < R)
            nd.push_back(y_cod ? leg.conj() : leg);
        else
            nc.push_back(y_cod ? leg : leg.conj());
    }
    return TensorStructure::create(y.sym, nd, nc);
}

PermutePlan make_permute_adj_plan(const StructPtr& xs, const TensorStructure& y, int shift, const std::vector<int>& p)
{
    const TensorStructure& in = *xs;
    check_perm(in, shift, p);
    const int R = in.rank, C = in.corank, N = R + C;
    const int newR = R - shift;
    const bool split = is_split_case(in, shift, p);
    {
        std::vector<Qbasis> nd, nc;
        permuted_bases(in, shift, p, split, nd, nc);
        StructPtr fwd = TensorStructure::create(in.sym, nd, nc);
        if(fwd->sector != y.sector || fwd->size != y.size) throw std::invalid_argument("permute_adj: the tensor does not have the permuted structure");
    }
    const Permutation P(p);
    std::unique_ptr<Permutation> Pd, Pc;
    if(split) {
        std::vector<int> pdv(p.begin(), p.begin() + R), pcv;
        for(int i = R; i < N; ++i) pcv.push_back(p[i] - R);
        Pd = std::make_unique<Permutation>(pdv);
        Pc = std::make_unique<Permutation>(pcv);
    }
    for(const Qbasis* b : {&y.dom(), &y.cod()})
        if(!b->data.empty() && !b->data[0].trees.empty()) (void)b->left_offset(b->data[0].trees[0]);
    PermutePlan plan;
    plan.nlegs = N;
    plan.new_rank = R;
    plan.out = xs;
    for(std::size_t i = 0; i < in.sector.size(); ++i) {
        const Q& q = in.sector[i];
        std::size_t roff = 0;
        for(const auto& t1 : in.dom().tree(q)) {
            std::size_t coff = 0;
            for(const auto& t2 : in.cod().tree(q)) {
                PermuteTask task;
                task.dst_off = static_cast<std::int64_t>(in.offsets[i] + roff + coff * in.rows[i]);
                task.dst_ld = static_cast<std::int64_t>(in.rows[i]);
                task.rows = static_cast<std::int64_t>(t1.dim);
                task.cols = static_cast<std::int64_t>(t2.dim);
                for(int m = 0; m < N; ++m) task.ext[m] = m < R ? t1.dims[m] : t2.dims[m - R];
                // X leg m is Y leg k = pi_inv[m]; its stride inside a Y subblock (rows: Y domain legs, cols: Y codomain legs)
                for(int m = 0; m < N; ++m) {
                    const int k = P.pi_inv[m];
                    std::int64_t st = 1;
                    if(k < newR) {
                        for(int j = 0; j < k; ++j) st *= task.ext[p[j]];
                        task.is_col[m] = 0;
                    } else {
                        for(int j = newR; j < k; ++j) st *= task.ext[p[j]];
                        task.is_col[m] = 1;
                    }
                    task.sstride[m] = st;
                }
                task.src_begin = static_cast<int>(plan.sources.size());
                auto add = [&](const FTree& a, const FTree& b, double c) {
                    if(std::abs(c) < PERMUTE_SKIP) return;
                    const int so = y.find_sector(a.coupled);
                    if(so < 0) throw std::logic_error("permute_adj: image sector missing");
                    const std::size_t ro = y.dom().left_offset(a), co = y.cod().left_offset(b);
                    plan.sources.push_back(PermuteSource{static_cast<std::int64_t>(y.offsets[so] + ro + co * y.rows[so]), static_cast<std::int64_t>(y.rows[so]), c});
                };
                if(split) {
                    TreeMap pd = tree_permute(in.sym, t1, *Pd);
                    TreeMap pc = tree_permute(in.sym, t2, *Pc);
                    for(const auto& [a, ca] : pd)
                        for(const auto& [b, cb] : pc) add(a, b, ca * cb);
                } else {
                    for(const auto& [pr, c] : pair_permute(in.sym, shift, t1, t2, P)) add(pr.first, pr.second, c);
                }
                task.nsrc = static_cast<int>(plan.sources.size()) - task.src_begin;
                plan.tasks.push_back(task);
                plan.moved_elems += t1.dim * t2.dim;
                coff += t2.dim;
            }
            roff += t1.dim;
        }
    }
    return plan;
}

ContractPerms get_permutations(const std::vector<int>& a1, int rank1, const std::vector<int>& a2, int rank2, int rankres)
{
    const int r1 = static_cast<int>(a1.size()), r2 = static_cast<int>(a2.size());
    int max1 = 0, max2 = 0;
    for(int x : a1) max1 = std::max(max1, x);
    for(int x : a2) max2 = std::max(max2, x);
    if(max1 != max2) throw std::invalid_argument("contract: the two index lists contract a different number of legs");
    ContractPerms out;
    out.p2.assign(r2, 0);
    std::vector<int> found2(r2, -1);
    for(int ind = 0; ind < max2; ++ind) {
        int val = static_cast<int>(std::find(a2.begin(), a2.end(), ind + 1) - a2.begin());
        found2[ind] = val;
        out.p2[ind] = val;
    }
    for(int ind = max2; ind < r2; ++ind) {
        int val = 0;
        for(int i = 0; i < r2; ++i)
            if(std::find(found2.begin(), found2.end(), i) == found2.end()) {
                val = i;
                break;
            }
        found2[ind] = val;
        out.p2[ind] = val;
    }
    out.p1.assign(r1, 0);
    std::vector<int> found1(r1, -1);
    for(int ind = 0; ind < max1; ++ind) {
        int val = static_cast<int>(std::find(a1.begin(), a1.end(), ind + 1) - a1.begin());
        found1[ind] = val;
        out.p1[r1 - max1 + ind] = val;
    }
    for(int ind = max1; ind < r1; ++ind) {
        int val = 0;
        for(int i = 0; i < r1; ++i)
            if(std::find(found1.begin(), found1.end(), i) == found1.end()) {
                val = i;
                break;
            }
        found1[ind] = val;
        out.p1[ind - max1] = val;
    }
    std::vector<int> tmp;
    for(int x : a1)
        if(x < 0) tmp.push_back(x);
    for(int x : a2)
        if(x < 0) tmp.push_back(x);
    out.pres.resize(tmp.size());
    std::iota(out.pres.begin(), out.pres.end(), 0);
    std::sort(out.pres.begin(), out.pres.end(), [&](int i, int j) { return tmp[i] > tmp[j]; });
    out.shift1 = rank1 - r1 + max1;
    out.shift2 = rank2 - max2;
    out.shiftres = r1 - max1 - rankres;
    return out;
}

GemmPlan make_gemm_plan(const TensorStructure& l, const TensorStructure& r)
{
    if(!(l.sym == r.sym)) throw std::invalid_argument("operator*: symmetry mismatch");
    if(l.corank != r.rank) throw std::invalid_argument("operator*: CoRank of the left tensor must equal Rank of the right tensor");
    GemmPlan plan;
    plan.out = TensorStructure::create(l.sym, l.unc_dom, r.unc_cod);
    const TensorStructure& o = *plan.out;
    for(std::size_t i = 0; i < l.sector.size(); ++i) {
        int j = r.find_sector(l.sector[i]);
        if(j < 0) continue;
        int so = o.find_sector(l.sector[i]);
        if(so < 0) throw std::logic_error("operator*: output sector missing");
        if(l.cols[i] != r.rows[j]) throw std::invalid_argument("operator*: inner block dimensions differ (codomain of left != domain of right)");
        GemmItem it;
        it.a_off = static_cast<std::int64_t>(l.offsets[i]);
        it.b_off = static_cast<std::int64_t>(r.offsets[j]);
        it.c_off = static_cast<std::int64_t>(o.offsets[so]);
        it.m = static_cast<int>(l.rows[i]);
        it.n = static_cast<int>(r.cols[j]);
        it.k = static_cast<int>(l.cols[i]);
        if(it.m > 0 && it.n > 0) {
            plan.items.push_back(it);
            plan.flops += 2.0 * it.m * it.n * static_cast<double>(it.k);
        }
    }
    return plan;
}

std::string q_str(const Sym& s, const Q& q)
{
    std::ostringstream os;
    os << "(";
    for(int c = 0; c < s.nq; ++c) os << (c ? "," : "") << q.v[c];
    os << ")";
    return os.str();
}

}  // namespace xa
