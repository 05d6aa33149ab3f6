// Prints draws of std::mt19937 + std::uniform_real_distribution<double>(-1,1) and
// std::normal_distribution<double>(0,1) from libstdc++ as exact hex floats.  Built and run by
// tests/test_oracle_pins.py to pin the oracle's restatement of the reference's generator.
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char** argv)
{
    unsigned seed = argc > 1 ? static_cast<unsigned>(std::strtoul(argv[1], nullptr, 10)) : 5489u;
    int n = argc > 2 ? std::atoi(argv[2]) : 16;
    std::mt19937 e(seed);
    std::uniform_real_distribution<double> u(-1., 1.);
    for(int i = 0; i < n; ++i) std::printf("%a\n", u(e));
    std::mt19937 e2(seed);
    std::normal_distribution<double> g(0., 1.);
    for(int i = 0; i < n; ++i) std::printf("%a\n", g(e2));
    return 0;
}
