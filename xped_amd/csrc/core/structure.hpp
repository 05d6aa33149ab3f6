// Host-side integer structure of block-sparse symmetric tensors: fusion trees, bases with
// fusion history, tree-pair recoupling, and the storage layout of a tensor.
//
// Semantics follow the reference's Core layer line by line where ordering is observable:
//   FusionTree          include/Xped/Core/FusionTree.hpp:22-115, src/Core/FusionTree.cpp
//   Qbasis              include/Xped/Core/Qbasis.hpp:37-268, src/Core/Qbasis.cpp
//   build_FusionTree    src/Core/TensorHelper.cpp:14-44
//   treepair            src/Core/treepair.cpp:35-217
//   StorageType layout  include/Xped/Core/storage/StorageType_contiguous.hpp:79-97,161-171
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "sym.hpp"

namespace xa {

constexpr int MAXR = 10;                 // maximal number of legs on one side of a tensor
constexpr double TREE_EPS = 1.e-12;      // ScalarTraits<double>::epsilon() (ScalarTraits.hpp:17)
constexpr double PERMUTE_SKIP = 1.e-10;  // |coeff| threshold in permute_impl (Tensor.cpp:127,193)

struct FTree {
    int rank = 0;
    std::array<Q, MAXR> unc{};
    std::array<Q, MAXR> inter{};  // rank-2 entries are meaningful (util::inter_dim)
    Q coupled{};
    std::array<std::uint32_t, MAXR> dims{};
    std::uint64_t dim = 1;
    std::array<bool, MAXR> dual{};

    int n_inter() const { return rank >= 2 ? rank - 2 : 0; }
    void compute_dim()
    {
        dim = 1;
        for(int i = 0; i < rank; ++i) dim *= dims[i];
    }
    // FusionTree::operator== without XPED_PEDANTIC_ASSERTS (FusionTree.cpp:66-83): IS_DUAL is ignored
    bool same(const FTree& o) const
    {
        if(rank != o.rank || coupled != o.coupled) return false;
        for(int i = 0; i < rank; ++i)
            if(unc[i] != o.unc[i] || dims[i] != o.dims[i]) return false;
        for(int i = 0; i < n_inter(); ++i)
            if(inter[i] != o.inter[i]) return false;
        return true;
    }
    // FusionTree::operator< (FusionTree.cpp:32-64): uncoupled, intermediates, coupled (multiplicities are all 0)
    bool less(const FTree& o) const
    {
        for(int i = 0; i < rank; ++i) {
            if(unc[i] < o.unc[i]) return true;
            if(unc[i] > o.unc[i]) return false;
        }
        for(int i = 0; i < n_inter(); ++i) {
            if(inter[i] < o.inter[i]) return true;
            if(inter[i] > o.inter[i]) return false;
        }
        return coupled < o.coupled;
    }
    std::uint64_t hash() const;               // consistent with same()
    FTree enlarge(const FTree& other) const;  // FusionTree.cpp:309-324
    void compute_intermediates(const Sym& s);  // FusionTree.cpp:136-152 (abelian only)
};

// Total order over every field (used as a map key inside recoupling; no reference counterpart
// for the order itself, which only influences floating-point accumulation order of SU(2) coefficients).
struct FTreeKeyLess {
    bool operator()(const FTree& a, const FTree& b) const;
};
using TreeMap = std::map<FTree, double, FTreeKeyLess>;
using TreePair = std::pair<FTree, FTree>;
struct TreePairLess {
    bool operator()(const TreePair& a, const TreePair& b) const
    {
        FTreeKeyLess l;
        if(l(a.first, b.first)) return true;
        if(l(b.first, a.first)) return false;
        return l(a.second, b.second);
    }
};
using PairMap = std::map<TreePair, double, TreePairLess>;

struct Permutation {
    std::vector<int> pi, pi_inv;
    explicit Permutation(std::vector<int> p);
    std::vector<int> decompose() const;  // Permutations.hpp:195-225
    bool is_identity() const;
};

// One basis entry (q, num, plain dim) plus its fusion trees.
struct QEntry {
    Q q;
    std::size_t num = 0;
    std::size_t dim = 0;
    std::vector<FTree> trees;
};

struct Qbasis {
    Sym sym;
    int depth = 1;
    bool conj_flag = false;
    bool sorted = false;
    std::size_t curr_dim = 0;
    std::vector<QEntry> data;
    // per sector: tree hash -> (tree index, left offset); built on the first left_offset call and
    // dropped by every mutation (push_back, push_back_trees, sort)
    using TreeIndex = std::vector<std::unordered_multimap<std::uint64_t, std::pair<std::size_t, std::size_t>>>;
    mutable std::shared_ptr<const TreeIndex> tree_index;

    Qbasis() = default;
    Qbasis(const Sym& s, int d) : sym(s), depth(d) {}

    int find(const Q& q) const
    {
        for(std::size_t i = 0; i < data.size(); ++i)
            if(data[i].q == q) return static_cast<int>(i);
        return -1;
    }
    bool present(const Q& q) const { return find(q) >= 0; }
    std::size_t inner_dim(const Q& q) const
    {
        int i = find(q);
        return i < 0 ? 0 : data[i].dim;
    }
    std::size_t dim() const
    {
        std::size_t d = 0;
        for(auto& e : data) d += e.dim;
        return d;
    }
    std::size_t full_dim() const
    {
        std::size_t d = 0;
        for(auto& e : data) d += e.dim * sym.degeneracy(e.q);
        return d;
    }
    const std::vector<FTree>& tree(const Q& q) const
    {
        int i = find(q);
        if(i < 0) throw std::runtime_error("Qbasis::tree: quantum number not present");
        return data[i].trees;
    }

    void push_back(const Q& q, std::size_t inner);                                    // Qbasis.cpp:34-73
    void push_back_trees(const Q& q, std::size_t inner, const std::vector<FTree>& t);  // Qbasis.cpp:75-91
    std::vector<std::size_t> sort();                                                    // Qbasis.cpp:306-331
    Qbasis conj() const;                                                                // Qbasis.cpp:418-439
    Qbasis combine(const Qbasis& other) const;                                          // Qbasis.cpp:441-495
    Qbasis forget_history() const;                                                      // Qbasis.cpp:410-416
    std::size_t left_offset(const FTree& t) const;                                      // Qbasis.cpp:208-225
    std::pair<std::size_t, std::size_t> tree_pos(const FTree& t) const;                 // (index in its sector's tree list, left offset)
    bool equals(const Qbasis& o) const;                                                 // Qbasis.cpp:333-349
};

Qbasis build_fusion_tree(const Sym& s, const std::vector<Qbasis>& uncoupled);  // TensorHelper.cpp:14-44
// build_fusion_tree through a process-wide cache keyed by the exact legs (charges, dimensions and
// dual flags of depth-1 legs); the result is immutable and shared
std::shared_ptr<const Qbasis> coupled_basis(const Sym& s, const std::vector<Qbasis>& uncoupled);

// Fusion tree moves.
TreeMap tree_swap(const Sym& s, const FTree& t, int pos);                       // FusionTree.cpp:353-412
TreeMap tree_permute(const Sym& s, const FTree& t, const Permutation& p);       // FusionTree.cpp:326-351
PairMap pair_turn_right(const Sym& s, const FTree& t1, const FTree& t2);        // treepair.cpp:35-88
PairMap pair_turn_left(const Sym& s, const FTree& t1, const FTree& t2);         // treepair.cpp:90-101
PairMap pair_turn(const Sym& s, int shift, const FTree& t1, const FTree& t2);   // treepair.cpp:103-141
PairMap pair_permute(const Sym& s, int shift, const FTree& t1, const FTree& t2, const Permutation& p);  // treepair.cpp:143-217

// Structure (no data) of a Tensor<Scalar, Rank, CoRank, Symmetry>.
struct TensorStructure {
    Sym sym;
    int rank = 0, corank = 0;
    std::vector<Qbasis> unc_dom, unc_cod;  // depth-1 legs
    // coupled (fusion-tree) bases, shared by all structures built from the same legs
    std::shared_ptr<const Qbasis> dom_p, cod_p;
    const Qbasis& dom() const { return *dom_p; }
    const Qbasis& cod() const { return *cod_p; }
    std::vector<Q> sector;
    std::vector<std::size_t> offsets, rows, cols;
    std::size_t size = 0;

    // Tensor(basis_domain, basis_codomain) + StorageType::resize (StorageType_contiguous.hpp:34-44,79-97)
    static std::shared_ptr<TensorStructure> create(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c);
    // Layout built by repeated StorageType::push_back (insertion order, StorageType_contiguous.hpp:161-171)
    static std::shared_ptr<TensorStructure> create_pushed(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c,
                                                          const std::vector<Q>& sectors);
    int find_sector(const Q& q) const
    {
        for(std::size_t i = 0; i < sector.size(); ++i)
            if(sector[i] == q) return static_cast<int>(i);
        return -1;
    }
    std::size_t leg_dim(int leg, const Q& q) const
    {
        return leg < rank ? unc_dom[leg].inner_dim(q) : unc_cod[leg - rank].inner_dim(q);
    }
};
using StructPtr = std::shared_ptr<TensorStructure>;

// Gather plan of one block-sparse permute: every output subblock lists its (source subblock, coeff)
// contributions in the order the reference accumulates them (Tensor.cpp:179-212).
struct PermuteSource {
    std::int64_t src_off;  // element offset of the source subblock's (0,0) in the input buffer
    std::int64_t src_ld;   // leading dimension of the source block
    double coeff;
};
// The builders write every scalar field and the first nlegs entries of each array; the empty
// constructor leaves the rest unwritten, so resizing a recycled table touches no memory.
struct PermuteTask {
    PermuteTask() {}
    std::int64_t dst_off;  // element offset of the output subblock's (0,0)
    std::int64_t dst_ld;
    std::int64_t rows, cols;  // output subblock extent
    int nsrc;
    int src_begin;
    std::array<std::uint32_t, MAXR * 2> ext;      // output leg extents
    std::array<std::int64_t, MAXR * 2> sstride;  // per output leg: stride in source row (is_col=0) or col (is_col=1) index
    std::array<std::uint8_t, MAXR * 2> is_col;
};
struct PermutePlan {
    int nlegs = 0, new_rank = 0;
    std::vector<PermuteTask> tasks;
    std::vector<PermuteSource> sources;
    StructPtr out;
    std::size_t moved_elems = 0;  // sum over tasks of rows*cols
};

// Dimension-free part of a permute plan: which input subblocks (sector, row tree, column tree),
// with which coefficients, feed every output subblock.  Tree lists are enumerated from charges
// alone, so this part serves every input structure with the same charges on every leg; only the
// offsets and extents (numeric_permute_plan) depend on the dimensions.
struct SymbolicPermute {
    struct Src {
        std::int32_t sector, t1, t2;
        double coeff;
    };
    int nlegs = 0, new_rank = 0;
    std::vector<std::int32_t> task_base, ncod;  // per output sector: first task, codomain trees
    std::vector<std::int32_t> src_begin, nsrc;  // per task (sector-major, then row tree, then column tree)
    std::vector<Src> sources;                   // grouped by task, each group in the reference's order
};
SymbolicPermute make_symbolic_permute(const TensorStructure& in, const TensorStructure& out, int shift, const std::vector<int>& p);
PermutePlan numeric_permute_plan(const SymbolicPermute& sp, const TensorStructure& in, const StructPtr& out, const std::vector<int>& p);

// Build the output structure and gather plan of T.permute<shift, p...>() (Tensor.cpp:83-214,335-355).
PermutePlan make_permute_plan(const TensorStructure& in, int shift, const std::vector<int>& p);
// Hand a plan's task and source tables back to this thread's buffers once the device copy exists:
// the next plan reuses their memory instead of faulting in fresh pages.  The plan keeps its
// output structure and counts; its tables are empty afterwards.
void recycle_plan_tables(PermutePlan& plan);
// Output structure of permute without the plan.
StructPtr permuted_structure(const TensorStructure& in, int shift, const std::vector<int>& p);
// Structure of X for a given Y = X.permute<shift, p...>() (the legs of Tensor::permute_adj,
// Tensor.cpp:270-288).
StructPtr permute_adj_structure(const TensorStructure& y, int shift, const std::vector<int>& p);
// Gather plan of Y.permute_adj<shift, p...>() (Tensor.cpp:258-333): the transpose of the permute
// map X -> Y.  Output = the structure `x` (which must permute to y), one task per subblock of x
// with the forward images of its tree pair as sources.
PermutePlan make_permute_adj_plan(const StructPtr& x, const TensorStructure& y, int shift, const std::vector<int>& p);

// get_permutations (Constfct.hpp:55-109) evaluated at run time.
struct ContractPerms {
    std::vector<int> p1, p2, pres;
    int shift1 = 0, shift2 = 0, shiftres = 0;
};
ContractPerms get_permutations(const std::vector<int>& a1, int rank1, const std::vector<int>& a2, int rank2, int rankres);

// Matched-sector GEMM list of operator* (TensorBase.cpp:237-288).
struct GemmItem {
    std::int64_t a_off, b_off, c_off;
    int m, n, k;
};
struct GemmPlan {
    std::vector<GemmItem> items;
    StructPtr out;
    double flops = 0;
};
GemmPlan make_gemm_plan(const TensorStructure& l, const TensorStructure& r);

std::string q_str(const Sym& s, const Q& q);

}  // namespace xa
