// Compiles the C++ facade (include/xped_amd.hpp) against the C ABI headers and runs it.
//   adapter_test struct  -- structure-only runtime: bases, combine, tensor sectors, error mapping
//   adapter_test gpu     -- device runtime: setRandom, contract, permute, tSVD with SVspec, tQR
// Prints one JSON object that the Python tests compare with the oracle.
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "xped_amd.hpp"

using namespace xped_amd;

static void print_sectors(const char* name, const Tensor& t, bool comma = true)
{
    std::printf("\"%s\": [", name);
    auto s = t.sectors();
    for(std::size_t i = 0; i < s.size(); ++i) std::printf("%s[%d, %zu, %zu, %zu]", i ? ", " : "", s[i].q[0], s[i].offset, s[i].rows, s[i].cols);
    std::printf("]%s\n", comma ? "," : "");
}

static void print_data(const char* name, const Tensor& t, bool comma = true)
{
    auto d = t.data();
    std::printf("\"%s\": [", name);
    for(std::size_t i = 0; i < d.size(); ++i) std::printf("%s%.17g", i ? ", " : "", d[i]);
    std::printf("]%s\n", comma ? "," : "");
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    try {
        Runtime rt(gpu ? 0 : -1);
        Basis virt(1), phys(1);
        for(int q = -2; q <= 2; ++q) virt.push_back({q}, 3 - (q < 0 ? -q : q));
        phys.push_back({-1}, 1).push_back({1}, 1);
        Basis fused = virt.combine(phys);
        std::printf("{\n\"virt_dim\": %zu, \"fused_dim\": %zu, \"fused_full\": %zu,\n", virt.dim(), fused.dim(), fused.fullDim());
        std::printf("\"fused\": [");
        auto e = fused.entries();
        for(std::size_t i = 0; i < e.size(); ++i) std::printf("%s[%d, %zu, %zu]", i ? ", " : "", e[i].first.first, e[i].second.first, e[i].second.second);
        std::printf("],\n");
        Tensor A(rt, {&virt, &phys}, {&virt}), B(rt, {&virt, &phys}, {&virt});
        print_sectors("A", A);
        if(!gpu) {
            bool mapped = false;
            try {
                A.data();  // no device: the status becomes std::runtime_error
            } catch(const std::runtime_error&) {
                mapped = true;
            }
            std::printf("\"error_mapped\": %s\n}\n", mapped ? "true" : "false");
            return 0;
        }
        A.setRandom(1);
        B.setRandom(2);
        print_data("A_data", A);
        print_data("B_data", B);
        // A(l,p|r) B(r,p'|r') -> (l,p|p',r'), then tSVD and tQR
        Tensor C = A.contract(B, {-1, -2, 1}, {1, -3, -4}, 2);
        print_sectors("C", C);
        print_data("C_data", C);
        Tensor P = C.permute(-1, {0, 3, 1, 2});
        print_sectors("P", P);
        print_data("P_data", P);
        double tw = 0.0, ent = 0.0;
        std::map<std::pair<int, int>, std::vector<double>> spec;
        auto [U, S, V] = C.tSVD(8, 0.0, tw, ent, spec, false, true);
        print_sectors("S", S);
        std::printf("\"truncWeight\": %.17g, \"entropy\": %.17g,\n\"SVspec\": {", tw, ent);
        bool first = true;
        for(const auto& [q, v] : spec) {
            std::printf("%s\"%d\": [", first ? "" : ", ", q.first);
            for(std::size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
            std::printf("]");
            first = false;
        }
        std::printf("},\n");
        auto [Q, R] = C.tQR();
        print_data("Q_data", Q);
        print_data("R_data", R, false);
        std::printf("}\n");
    } catch(const std::exception& ex) {
        std::fprintf(stderr, "adapter_test: %s\n", ex.what());
        return 1;
    }
    return 0;
}
