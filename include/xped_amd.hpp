// xped_amd.hpp -- header-only C++ facade over the C ABI of include/xped_tensor.h.
//
// The maintainer-side adapter that binds Xped's Tensor API to the MI355X path: RAII handles for
// the runtime, bases and tensors, the reference's operation names (permute<shift, p...>,
// contract<a1, a2, ResRank>, operator*, tSVD, tQR, teigh, setRandom), and its error mapping
// (StorageType_contiguous.hpp:91-96: allocation failure -> std::bad_alloc; every other status ->
// std::runtime_error with the runtime's message).  Xped's own Tensor<Scalar, Rank, CoRank,
// Symmetry> keeps one xped_amd::Tensor per instance and forwards its operations here.
#pragma once

#include <initializer_list>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "xped_tensor.h"

namespace xped_amd {

inline void check(int rc, const xt_rt* rt, const char* what)
{
    if(rc == XH_OK) return;
    if(rc == XH_ENOMEM) throw std::bad_alloc();
    const char* msg = xt_last_error(rt);
    throw std::runtime_error(std::string(what) + ": " + (msg ? msg : ""));
}

// Runtime: the device context and plan caches of one host thread (device < 0: structure only).
class Runtime {
public:
    explicit Runtime(int device = 0) { check(xt_rt_create(device, &rt_), nullptr, "xt_rt_create"); }
    ~Runtime()
    {
        if(rt_) xt_rt_destroy(rt_);
    }
    Runtime(const Runtime&) = delete;
    Runtime& operator=(const Runtime&) = delete;
    xt_rt* get() const { return rt_; }

private:
    xt_rt* rt_ = nullptr;
};

// Qbasis<Symmetry, 1>: kind0/kind1 as in xt_basis_create (1 U(1), 2 fermionic U(1), 3 SU(2)).
class Basis {
public:
    Basis(int kind0, int kind1 = 0) { check(xt_basis_create(kind0, kind1, &b_), nullptr, "xt_basis_create"); }
    explicit Basis(xt_basis* b) : b_(b) {}
    ~Basis()
    {
        if(b_) xt_basis_destroy(b_);
    }
    Basis(const Basis&) = delete;
    Basis& operator=(const Basis&) = delete;
    Basis(Basis&& o) noexcept : b_(std::exchange(o.b_, nullptr)) {}
    Basis& push_back(std::initializer_list<int> q, std::size_t dim)  // Qbasis.cpp:34-73
    {
        int qq[2] = {0, 0};
        int i = 0;
        for(int x : q) qq[i++ & 1] = x;
        check(xt_basis_push_back(b_, qq, dim), nullptr, "Qbasis::push_back");
        return *this;
    }
    Basis combine(const Basis& o) const  // Qbasis.cpp:441-495
    {
        xt_basis* out = nullptr;
        check(xt_basis_combine(b_, o.b_, &out), nullptr, "Qbasis::combine");
        return Basis(out);
    }
    Basis conj() const
    {
        xt_basis* out = nullptr;
        check(xt_basis_conj(b_, &out), nullptr, "Qbasis::conj");
        return Basis(out);
    }
    std::size_t dim() const
    {
        int depth, nsec, conj;
        std::size_t d, fd;
        check(xt_basis_info(b_, &depth, &nsec, &d, &fd, &conj), nullptr, "Qbasis::dim");
        return d;
    }
    std::size_t fullDim() const
    {
        int depth, nsec, conj;
        std::size_t d, fd;
        check(xt_basis_info(b_, &depth, &nsec, &d, &fd, &conj), nullptr, "Qbasis::fullDim");
        return fd;
    }
    // (q, num, inner dim) of every sector, in basis order
    std::vector<std::pair<std::pair<int, int>, std::pair<std::size_t, std::size_t>>> entries() const
    {
        int depth, nsec, conj;
        std::size_t d, fd;
        check(xt_basis_info(b_, &depth, &nsec, &d, &fd, &conj), nullptr, "Qbasis entries");
        std::vector<std::pair<std::pair<int, int>, std::pair<std::size_t, std::size_t>>> out;
        for(int i = 0; i < nsec; ++i) {
            int q[2], nt;
            std::size_t num, dim;
            check(xt_basis_entry(b_, i, q, &num, &dim, &nt), nullptr, "Qbasis entry");
            out.push_back({{q[0], q[1]}, {num, dim}});
        }
        return out;
    }
    const xt_basis* get() const { return b_; }

private:
    xt_basis* b_ = nullptr;
};

struct Sector {
    int q[2];
    std::size_t offset, rows, cols;
};

// Tensor<double, Rank, CoRank, Symmetry> with its data on the device.
class Tensor {
public:
    Tensor(const Runtime& rt, const std::vector<const Basis*>& domain, const std::vector<const Basis*>& codomain) : rt_(rt.get())
    {
        std::vector<const xt_basis*> legs;
        for(auto* b : domain) legs.push_back(b->get());
        for(auto* b : codomain) legs.push_back(b->get());
        check(xt_tensor_create(rt_, int(domain.size()), int(codomain.size()), legs.data(), &t_), rt_, "Tensor()");
    }
    Tensor(xt_rt* rt, xt_tensor* t) : rt_(rt), t_(t) {}
    ~Tensor()
    {
        if(t_) xt_tensor_destroy(t_);
    }
    Tensor(const Tensor&) = delete;
    Tensor& operator=(const Tensor&) = delete;
    Tensor(Tensor&& o) noexcept : rt_(o.rt_), t_(std::exchange(o.t_, nullptr)) {}

    std::vector<Sector> sectors() const
    {
        int r, c, n;
        std::size_t size;
        check(xt_tensor_shape(t_, &r, &c, &n, &size), rt_, "sectors");
        std::vector<Sector> out(n);
        for(int i = 0; i < n; ++i) check(xt_tensor_sector(t_, i, out[i].q, &out[i].offset, &out[i].rows, &out[i].cols), rt_, "sector");
        return out;
    }
    std::size_t size() const
    {
        int r, c, n;
        std::size_t size;
        check(xt_tensor_shape(t_, &r, &c, &n, &size), rt_, "size");
        return size;
    }
    Tensor& setRandom(unsigned seed)  // Tensor.cpp:32-43 with std::mt19937(seed)
    {
        check(xt_tensor_set_random(t_, seed, 0), rt_, "setRandom");
        return *this;
    }
    std::vector<double> data() const
    {
        std::vector<double> h(size());
        check(xt_tensor_download(t_, h.data(), h.size()), rt_, "download");
        return h;
    }
    Tensor permute(int shift, std::initializer_list<int> p) const  // Tensor.hpp:297-316
    {
        xt_tensor* out = nullptr;
        std::vector<int> pv(p);
        check(xt_permute(t_, shift, pv.data(), int(pv.size()), &out), rt_, "permute");
        return Tensor(rt_, out);
    }
    Tensor operator*(const Tensor& o) const  // TensorBase.cpp:237-288
    {
        xt_tensor* out = nullptr;
        check(xt_mul(t_, o.t_, XH_FP64, &out), rt_, "operator*");
        return Tensor(rt_, out);
    }
    Tensor contract(const Tensor& o, std::initializer_list<int> a1, std::initializer_list<int> a2, int rankres) const  // Tensor.hpp:328-351
    {
        xt_tensor* out = nullptr;
        std::vector<int> v1(a1), v2(a2);
        check(xt_contract(t_, v1.data(), int(v1.size()), o.t_, v2.data(), int(v2.size()), rankres, XH_FP64, &out), rt_, "contract");
        return Tensor(rt_, out);
    }
    // tSVD(maxKeep, eps_svd, truncWeight, entropy, SVspec, PRESERVE_MULTIPLETS, RETURN_SPEC) (Tensor.hpp:353-362)
    std::tuple<Tensor, Tensor, Tensor> tSVD(std::size_t maxKeep, double eps_svd, double& truncWeight, double& entropy, std::map<std::pair<int, int>, std::vector<double>>& SVspec,
             bool PRESERVE_MULTIPLETS = true, bool RETURN_SPEC = true) const
    {
        xt_spec* sp = nullptr;
        check(xt_spec_create(&sp), rt_, "tSVD");
        std::unique_ptr<xt_spec, int (*)(xt_spec*)> guard(sp, xt_spec_destroy);
        xt_tensor *u = nullptr, *s = nullptr, *v = nullptr;
        check(xt_tsvd_spec(t_, maxKeep, eps_svd, &truncWeight, &entropy, sp, PRESERVE_MULTIPLETS, RETURN_SPEC, &u, &s, &v), rt_, "tSVD");
        int n = 0;
        xt_spec_size(sp, &n);
        for(int i = 0; i < n; ++i) {
            int q[2], cnt = 0;
            xt_spec_entry(sp, i, q, &cnt, nullptr, 0);
            std::vector<double> vals(cnt);
            xt_spec_entry(sp, i, q, &cnt, vals.data(), cnt);
            SVspec.insert({{q[0], q[1]}, vals});
        }
        return {Tensor(rt_, u), Tensor(rt_, s), Tensor(rt_, v)};
    }
    std::pair<Tensor, Tensor> tQR(bool RETURN_LQ = false) const  // Tensor.hpp:390-391
    {
        xt_tensor *q = nullptr, *r = nullptr;
        check(xt_tqr(t_, RETURN_LQ, &q, &r), rt_, "tQR");
        return {Tensor(rt_, q), Tensor(rt_, r)};
    }
    std::pair<Tensor, Tensor> teigh() const  // Tensor.hpp:393-394
    {
        xt_tensor *d = nullptr, *v = nullptr;
        check(xt_teigh(t_, &d, &v), rt_, "teigh");
        return {Tensor(rt_, d), Tensor(rt_, v)};
    }
    const xt_tensor* get() const { return t_; }

private:
    xt_rt* rt_ = nullptr;
    xt_tensor* t_ = nullptr;
};

}  // namespace xped_amd
