// Device-resident block-sparse tensor and the operations of the reference's Xped::Tensor
// (include/Xped/Core/Tensor.hpp:59-444, TensorBase.hpp:47-135) on top of the HIP device layer.
#pragma once
#include <array>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/xped_hip.h"
#include "structure.hpp"

namespace xa {

// Error carrying an xh status code (XH_EINVAL, XH_ENOMEM, XH_EHIP, XH_ENOTCONV, XH_ENODEV).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Owns a device context; buffers and the runtime keep it alive until the last one is gone.
struct CtxHolder {
    xh_ctx* ctx = nullptr;
    ~CtxHolder()
    {
        if(ctx) xh_destroy(ctx);
    }
};

struct DeviceBuffer {
    std::shared_ptr<CtxHolder> holder;
    double* ptr = nullptr;
    std::size_t n = 0;
    ~DeviceBuffer()
    {
        if(holder && holder->ctx && ptr) xh_free(holder->ctx, ptr);
    }
};

struct CachedPermute {
    PermutePlan plan;
    xh_permute_plan* dev = nullptr;
    std::uint64_t last_use = 0;  // runtime tick of the last lookup (cache eviction order)
};
struct CachedGemm {
    GemmPlan plan;
    xh_gemm_plan* dev = nullptr;
    std::vector<xh_copy_rect> zero_rects;  // output sectors not written by any GEMM
    std::vector<std::int64_t> zero_ranges;  // the same as (offset, length) pairs, adjacent sectors merged
    std::uint64_t last_use = 0;
};

// contract() with the permute of its left operand fused into the product (xh_gemm_gather): the
// permuted structure, the GEMM plan of (permuted left, permuted right) and the device gather plan;
// dev is null when the product does not qualify (then contract() permutes and multiplies)
struct CachedGather {
    StructPtr permuted;
    std::shared_ptr<CachedGemm> gemm;
    xh_gather_plan* dev = nullptr;
    std::uint64_t last_use = 0;
    // the product's own permute folded into the same launch (xh_gemm_gather_scatter), per
    // (shift, p) of that permute: the permuted product's structure and the device plan (null when
    // the permute cannot be folded; contract() then permutes the product)
    struct Scatter {
        StructPtr out;
        xh_scatter_plan* dev = nullptr;
    };
    std::unordered_map<std::string, Scatter> scatter;
    void release(xh_ctx* ctx);  // device plans, scatter plans first
};

struct Tape;

// Device context, plan caches and counters.  ctx == nullptr is a structure-only runtime: every
// integer/structure operation works, data operations fail with XH_ENODEV.
struct Runtime {
    std::shared_ptr<CtxHolder> holder;
    xh_ctx* ctx = nullptr;
    std::unordered_map<std::string, std::shared_ptr<CachedPermute>> permute_cache;
    // dimension-free permute plans keyed by the charges of the input legs, shift and p
    std::unordered_map<std::string, std::shared_ptr<const SymbolicPermute>> sym_permute_cache;
    long n_symbolic_hits = 0;
    std::unordered_map<std::string, std::shared_ptr<CachedGemm>> gemm_cache;
    std::unordered_map<std::string, std::shared_ptr<CachedGather>> gather_cache;
    // contract(): 1 (default) reads the left operand through its permute inside skinny FP64
    // products instead of materialising the permuted tensor; 2 also writes the product straight
    // into its own permuted layout; 0 materialises both.  All give the same results bit for bit.
    int fuse_permute = 1;
    long n_fused = 0;
    std::unordered_map<std::string, StructPtr> struct_cache;
    // counters for reporting
    double gemm_flops = 0.0, permute_bytes = 0.0;
    long n_gemm = 0, n_permute = 0;
    // plan caches are bounded: past max_plans entries the least recently used half is dropped
    // (device tables freed stream-ordered, so launches already queued still read them)
    std::uint64_t tick = 0;
    std::size_t max_plans = 2048;
    long n_evicted = 0;
    void trim_caches();
    // host time spent building plans (seconds) and the number built
    double permute_plan_s = 0.0, gemm_plan_s = 0.0;
    long n_permute_plans = 0, n_gemm_plans = 0;
    // breakdown (seconds): 0 host permute plan, 1 its checks and packing, 2 device permute plan,
    // 3 host GEMM plan, 4 device GEMM plan, 5 output structures of mul
    double host_t[8] = {};
    // optional per-launch timing of the contraction GEMMs (operator*) with HIP events, and a log
    // of their block shapes (m, n, k); off by default (it synchronises after every launch)
    bool time_gemm = false;
    std::vector<std::array<int32_t, 3>> gemm_log;
    // reverse-mode recording: while tape->active, every differentiable operation pushes its reverse pass
    Tape* tape = nullptr;
    ~Runtime();
    void require_device() const;
};

std::string structure_key(const TensorStructure& t);
std::string legs_key(const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c);
StructPtr cached_structure(Runtime* rt, const Sym& s, const std::vector<Qbasis>& d, const std::vector<Qbasis>& c);

class DTensor {
public:
    StructPtr st;
    std::shared_ptr<DeviceBuffer> buf;
    Runtime* rt = nullptr;

    DTensor() = default;
    double* data() const { return buf ? buf->ptr : nullptr; }
    std::size_t size() const { return st ? st->size : 0; }

    static DTensor empty(Runtime* rt, StructPtr st);  // uninitialised data
    static DTensor zeros(Runtime* rt, StructPtr st);  // setZero
    void upload(const double* host, std::size_t n);
    void download(double* host, std::size_t n) const;
    // setRandom (Tensor.cpp:32-43, MatrixInterface_Eigen_impl.cpp:51-64): one engine, blocks in
    // storage order, column-major within a block.  dist 0: uniform_real_distribution(-1,1),
    // dist 1: normal_distribution(0,1).
    void set_random(std::mt19937& engine, int dist);
};

// Reverse-mode recording of the tensor operations: the role of stan-math's reverse-pass callback
// stack in the reference's AD tensor (include/Xped/AD/ADTensor.hpp, vari_value.hpp:97-165).  While
// a Tape is set on a runtime and active, permute, operator*, adjoint, twist, scaled, axpby,
// diag_map, full_svd, cutoff (and so contract and tsvd), tqr and div_trace push their reverse
// passes; backward() runs them in reverse order.  Adjoints are keyed by the value's device buffer
// and accumulate (+=) as the reference's callbacks do.  Operations that return their input (an
// identity permutation, a bosonic twist) record nothing: value and adjoint are shared.
struct Tape {
    bool active = false;
    std::vector<std::function<void(Tape&)>> backs;
    struct Entry {
        DTensor value, adj;
    };
    std::unordered_map<const DeviceBuffer*, Entry> adj;

    const DTensor* find(const DTensor& x) const;  // nullptr when no adjoint reached x
    DTensor adjoint_of(const DTensor& x);         // created as zeros on first use; shares the stored buffer
    void accumulate(const DTensor& x, const DTensor& dx, double alpha = 1.0);  // adj(x) += alpha dx
    void backward();                              // stops recording, runs the reverse passes, drops them
    void clear();                                 // drops reverse passes and adjoints
};
inline Tape* recording(const DTensor& t)
{
    return (t.rt && t.rt->tape && t.rt->tape->active && t.buf) ? t.rt->tape : nullptr;
}
// x / trace(s) with s a square tensor (the weighted trace of TensorBase.cpp:26-37); recorded with
// the adjoint of both x and s.
DTensor div_trace(const DTensor& x, const DTensor& s);

DTensor permute(const DTensor& t, int shift, const std::vector<int>& p);
// Y.permute_adj<shift, p...>() (Tensor.cpp:258-333): the transpose of X -> X.permute<shift, p...>().
DTensor permute_adj(const DTensor& y, int shift, const std::vector<int>& p);
std::shared_ptr<CachedPermute> get_permute_plan(Runtime* rt, const TensorStructure& in, int shift, const std::vector<int>& p);
DTensor mul(const DTensor& l, const DTensor& r, int precision);
// operator* into an existing tensor with the product's structure: no allocation, so the launches
// can be captured in a graph once the plan exists
void mul_into(const DTensor& l, const DTensor& r, DTensor& out, int precision);
DTensor contract(const DTensor& a, const std::vector<int>& a1, const DTensor& b, const std::vector<int>& a2, int rankres, int precision);
// operator* of independent pairs in one grouped GEMM launch (each product as mul computes it)
std::vector<DTensor> mul_batch(const std::vector<std::pair<DTensor, DTensor>>& pairs, int precision);
struct ContractJob {
    DTensor a;
    std::vector<int> a1;
    DTensor b;
    std::vector<int> a2;
    int rank = 0;
};
// contract() of independent jobs with their products in one grouped GEMM launch
std::vector<DTensor> contract_batch(const std::vector<ContractJob>& jobs, int precision);
DTensor adjoint(const DTensor& t);
DTensor twist(const DTensor& t, int leg);
DTensor scaled(const DTensor& t, double alpha);
DTensor axpby(const DTensor& x, const DTensor& y, double a, double b);  // a*x + b*y, same structure
DTensor diag_map(const DTensor& t, int mode);                          // 0 sqrt, 1 inv, 2 inv sqrt of diagonal blocks
// Tensor::Identity(dom, cod) (Tensor.hpp:229-235)
DTensor identity(Runtime* rt, const std::vector<Qbasis>& dom, const std::vector<Qbasis>& cod);
// TensorBase::mexp(factor) (TensorBase.cpp:180-183): exp(factor * block) of every square block
// (order <= 64)
DTensor mexp(const DTensor& t, double factor);
double trace(const DTensor& t);                                         // degeneracy-weighted (TensorBase.cpp:26-37)
double squared_norm(const DTensor& t);                                  // (T * T^dagger).trace()
double max_norm(const DTensor& t);                                      // TensorBase.cpp:48-56

struct SVDResult {
    DTensor U, S, Vdag;
    double trunc_weight = 0.0, entropy = 0.0;
    std::map<Q, std::vector<double>> spec;  // kept singular values per sector (RETURN_SPEC)
    std::size_t n_states = 0;               // number of singular values before truncation
};
// Tensor::tSVD (Tensor.cpp:478-633).  trunc_weight and entropy are accumulated into the values
// passed in, as the reference does.
SVDResult tsvd(const DTensor& t, std::size_t max_keep, double eps_svd, bool preserve_multiplets, double trunc_weight_in = 0.0,
               double entropy_in = 0.0);
// full_svd_impl alone: untruncated U, Sigma, Vdag plus all singular values in sector order
struct FullSVD {
    DTensor U, S, Vdag;
    std::vector<std::pair<Q, double>> all_sv;
    std::vector<Q> sectors;
};
FullSVD full_svd(const DTensor& t);
SVDResult cutoff(const DTensor& t, const FullSVD& f, std::size_t max_keep, double eps_svd, bool preserve_multiplets, double trunc_weight_in,
                 double entropy_in);
// Tensor::tQR(RETURN_LQ) (Tensor.cpp:757-780).  With return_lq the reference returns the
// adjoints of the per-sector factors (Q^T, R^T: the LQ factors of A^T) in the tensors of Q and R,
// which is only well-formed when every sector block is square; other blocks raise XH_EINVAL.
std::pair<DTensor, DTensor> tqr(const DTensor& t, bool return_lq = false);
std::pair<DTensor, DTensor> teigh(const DTensor& t);  // Tensor.cpp:782-803

// Host-side truncation rule of cutoff_matrices (Tensor.cpp:523-633) on an explicit list of
// (sector, singular value) pairs.  Returns the kept count per sector (ascending q).
struct TruncRule {
    std::map<Q, std::vector<double>> kept;
    double trunc_weight = 0.0, entropy = 0.0;
};
TruncRule truncation_rule(const Sym& s, std::vector<std::pair<Q, double>> all_sv, std::size_t max_keep, double eps_svd, bool preserve_multiplets,
                          double trunc_weight_in, double entropy_in);

}  // namespace xa
