// Storage wire format of a block-sparse tensor.
//
// The reference serialises StorageType with yas (include/Xped/Core/storage/StorageType_contiguous.hpp:
// 191-204: data, dict, sectors, uncoupled domain and codomain, coupled domain and codomain,
// offsets) and writes Tensor::save to HDF5 (src/Core/Tensor.cpp:984-1000: per sector its q and
// its block).  Neither library is available here, so the same fields travel in one fixed binary
// layout (little-endian):
//
//   "XPEDTNS1"                                  magic
//   int32 kind0, kind1, rank, corank            symmetry and leg counts
//   per uncoupled leg (domain, then codomain):  uint8 conj, uint8 sorted, uint64 curr_dim,
//                                               uint32 n, then n x (int32 q0, q1, uint64 num, dim, uint8 dual)
//   uint64 nsec, then nsec x (int32 q0, q1, uint64 offset, rows, cols)
//   uint64 ndata, then ndata doubles            the blocks, column-major, in sector order
//
// The coupled bases and the sector dictionary are functions of the legs and the sector list, so
// the reader rebuilds them (build_FusionTree, StorageType::push_back order) and refuses an archive
// whose offsets or block shapes differ from the rebuilt ones.  A structure-only runtime writes
// ndata = 0 and reads the structure alone.
#include <cstring>
#include <fstream>

#include "storage.hpp"

namespace xa {

namespace {
struct Writer {
    std::vector<char> b;
    template <typename T>
    void put(const T& x)
    {
        const char* p = reinterpret_cast<const char*>(&x);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void put_q(const Q& q)
    {
        put<int32_t>(q.v[0]);
        put<int32_t>(q.v[1]);
    }
};
struct Reader {
    const char* p;
    std::size_t n, at = 0;
    template <typename T>
    T get()
    {
        if(at + sizeof(T) > n) throw Error(XH_EINVAL, "tensor archive: truncated");
        T x;
        std::memcpy(&x, p + at, sizeof(T));
        at += sizeof(T);
        return x;
    }
    Q get_q()
    {
        Q q;
        q.v[0] = get<int32_t>();
        q.v[1] = get<int32_t>();
        return q;
    }
};
constexpr char MAGIC[9] = "XPEDTNS1";

void put_leg(Writer& w, const Qbasis& b)
{
    if(b.depth != 1) throw Error(XH_EINVAL, "tensor archive: legs must be depth-1 bases");
    w.put<uint8_t>(b.conj_flag);
    w.put<uint8_t>(b.sorted);
    w.put<uint64_t>(b.curr_dim);
    w.put<uint32_t>(static_cast<uint32_t>(b.data.size()));
    for(const auto& e : b.data) {
        w.put_q(e.q);
        w.put<uint64_t>(e.num);
        w.put<uint64_t>(e.dim);
        w.put<uint8_t>(!e.trees.empty() && e.trees[0].dual[0]);
    }
}

Qbasis get_leg(Reader& r, const Sym& s)
{
    Qbasis b(s, 1);
    b.conj_flag = r.get<uint8_t>() != 0;
    b.sorted = r.get<uint8_t>() != 0;
    b.curr_dim = r.get<uint64_t>();
    const uint32_t n = r.get<uint32_t>();
    if(n > (1u << 24)) throw Error(XH_EINVAL, "tensor archive: implausible sector count");
    for(uint32_t i = 0; i < n; ++i) {
        QEntry e;
        e.q = r.get_q();
        e.num = r.get<uint64_t>();
        e.dim = r.get<uint64_t>();
        FTree t;
        t.rank = 1;
        t.unc[0] = e.q;
        t.coupled = e.q;
        if(e.dim > 0xffffffffull) throw Error(XH_EINVAL, "tensor archive: sector dimension too large");
        t.dims[0] = static_cast<uint32_t>(e.dim);
        t.dual[0] = r.get<uint8_t>() != 0;
        t.compute_dim();
        e.trees.push_back(t);
        b.data.push_back(std::move(e));
    }
    return b;
}
}  // namespace

std::vector<char> serialize(const DTensor& t)
{
    const TensorStructure& s = *t.st;
    Writer w;
    w.b.insert(w.b.end(), MAGIC, MAGIC + 8);
    w.put<int32_t>(s.sym.kind[0]);
    w.put<int32_t>(s.sym.kind[1]);
    w.put<int32_t>(s.rank);
    w.put<int32_t>(s.corank);
    for(const auto& b : s.unc_dom) put_leg(w, b);
    for(const auto& b : s.unc_cod) put_leg(w, b);
    w.put<uint64_t>(s.sector.size());
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        w.put_q(s.sector[i]);
        w.put<uint64_t>(s.offsets[i]);
        w.put<uint64_t>(s.rows[i]);
        w.put<uint64_t>(s.cols[i]);
    }
    const std::size_t nd = t.rt && t.rt->ctx ? s.size : 0;
    w.put<uint64_t>(nd);
    if(nd) {
        const std::size_t at = w.b.size();
        w.b.resize(at + nd * sizeof(double));
        std::vector<double> h(nd);
        t.download(h.data(), nd);
        std::memcpy(w.b.data() + at, h.data(), nd * sizeof(double));
    }
    return w.b;
}

DTensor deserialize(Runtime* rt, const char* buf, std::size_t n)
{
    Reader r{buf, n};
    if(n < 8 || std::memcmp(buf, MAGIC, 8) != 0) throw Error(XH_EINVAL, "tensor archive: bad magic (not an XPEDTNS1 archive)");
    r.at = 8;
    const int k0 = r.get<int32_t>(), k1 = r.get<int32_t>();
    const int rank = r.get<int32_t>(), corank = r.get<int32_t>();
    if(rank < 0 || corank < 0 || rank + corank > 16) throw Error(XH_EINVAL, "tensor archive: bad rank");
    const Sym s = Sym::make(k0, k1);
    std::vector<Qbasis> dom, cod;
    for(int i = 0; i < rank; ++i) dom.push_back(get_leg(r, s));
    for(int i = 0; i < corank; ++i) cod.push_back(get_leg(r, s));
    const uint64_t nsec = r.get<uint64_t>();
    if(nsec > (1ull << 24)) throw Error(XH_EINVAL, "tensor archive: implausible sector count");
    std::vector<Q> sectors(nsec);
    std::vector<uint64_t> off(nsec), rows(nsec), cols(nsec);
    for(uint64_t i = 0; i < nsec; ++i) {
        sectors[i] = r.get_q();
        off[i] = r.get<uint64_t>();
        rows[i] = r.get<uint64_t>();
        cols[i] = r.get<uint64_t>();
    }
    StructPtr st = TensorStructure::create_pushed(s, dom, cod, sectors);
    for(uint64_t i = 0; i < nsec; ++i)
        if(st->offsets[i] != off[i] || st->rows[i] != rows[i] || st->cols[i] != cols[i])
            throw Error(XH_EINVAL, "tensor archive: the stored offsets or block shapes differ from the structure its legs give");
    const uint64_t nd = r.get<uint64_t>();
    if(nd != 0 && nd != st->size) throw Error(XH_EINVAL, "tensor archive: data length does not match the structure");
    DTensor t = DTensor::zeros(rt, st);
    if(nd) {
        if(r.at + nd * sizeof(double) > n) throw Error(XH_EINVAL, "tensor archive: truncated data");
        std::vector<double> h(nd);
        std::memcpy(h.data(), buf + r.at, nd * sizeof(double));
        t.upload(h.data(), nd);
    }
    return t;
}

void save(const DTensor& t, const std::string& path)
{
    const auto b = serialize(t);
    std::ofstream f(path, std::ios::binary);
    if(!f) throw Error(XH_EINVAL, "tensor save: cannot open " + path);
    f.write(b.data(), static_cast<std::streamsize>(b.size()));
    if(!f) throw Error(XH_EINVAL, "tensor save: write failed for " + path);
}

DTensor load(Runtime* rt, const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if(!f) throw Error(XH_EINVAL, "tensor load: cannot open " + path);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return deserialize(rt, b.data(), b.size());
}

}  // namespace xa
