// Sorts (value, index) pairs read from stdin with std::sort ordered by value descending -- the
// comparator of cutoff_matrices -- and prints the resulting index order.  Built and run by
// tests/test_oracle_pins.py to pin the oracle's restatement of libstdc++'s introsort.
#include <algorithm>
#include <cstdio>
#include <utility>
#include <vector>

int main()
{
    int n;
    while(std::scanf("%d", &n) == 1) {
        std::vector<std::pair<int, double>> v(n);
        for(int i = 0; i < n; ++i) {
            std::scanf("%lf", &v[i].second);
            v[i].first = i;
        }
        std::sort(v.begin(), v.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.second > b.second; });
        for(int i = 0; i < n; ++i) std::printf("%d ", v[i].first);
        std::printf("\n");
    }
    return 0;
}
