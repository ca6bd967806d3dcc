// The batched rho's reduction passes (reduce_waves.hpp) against a brute-force restatement of
// their rule, on random span sets with shared starts, nesting, touching ends and empty spans.
#include <cstdio>
#include <random>
#include <vector>

#include "reduce_waves.hpp"

using cfgmm::RhoSpan;

static std::vector<int> brute(const std::vector<RhoSpan> &sp) {
    std::vector<int> w(sp.size(), 0);  // sp[i].index == i
    for (size_t i = 0; i < sp.size(); ++i)
        for (size_t j = 0; j < i; ++j)
            if (sp[i].lo < sp[i].hi && sp[j].lo < sp[j].hi && sp[i].lo < sp[j].hi && sp[j].lo < sp[i].hi)
                w[i] = std::max(w[i], w[j] + 1);
    return w;
}

int main() {
    std::mt19937_64 rng(7);
    int failed = 0, cases = 0;
    for (int t = 0; t < 3000; ++t) {
        const size_t n = 1 + rng() % (t < 2000 ? 12 : 200);
        const uintptr_t range = 1 + rng() % (t % 3 == 0 ? 8 : 1000);
        std::vector<RhoSpan> sp(n);
        for (size_t i = 0; i < n; ++i) {
            const uintptr_t lo = 4096 + rng() % range;
            const uintptr_t len = (rng() % 5 == 0) ? 0 : 1 + rng() % (t % 2 ? 4 : 300);
            sp[i] = RhoSpan{lo, lo + len, (int)i};
        }
        std::vector<int> got;
        const int nw = cfgmm::assign_reduce_waves_spans(sp, got);
        const std::vector<int> want = brute(sp);
        int want_nw = 1;
        for (int x : want) want_nw = std::max(want_nw, x + 1);
        ++cases;
        if (got != want || nw != want_nw) {
            std::printf("FAIL case %d (n=%zu): passes %d vs %d\n", t, n, nw, want_nw);
            ++failed;
        }
    }
    // disjoint spans (the common case) take one pass
    std::vector<RhoSpan> dis;
    for (int i = 0; i < 100; ++i) dis.push_back(RhoSpan{(uintptr_t)(1000 + 10 * (99 - i)), (uintptr_t)(1010 + 10 * (99 - i)), i});
    std::vector<int> w;
    ++cases;
    if (cfgmm::assign_reduce_waves_spans(dis, w) != 1) ++failed;
    std::printf("waves_checks: %d cases, %d failed\n", cases, failed);
    return failed ? 1 : 0;
}
