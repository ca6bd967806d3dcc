// reduce_waves.hpp -- host-only: the order in which a batched rho call adds its problems'
// partial sums into rho (rho_batched.hip), kept apart from device code so the CPU checks can
// compile it.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace cfgmm {

// The rho range [lo, hi) of a problem (bytes) and its position in the call, 0..n-1.
struct RhoSpan {
    uintptr_t lo, hi;
    int index;
};

// Reduction passes: problem i goes one pass after the last earlier problem (in call order) whose
// range overlaps its own, so problems sharing rho elements add in call order, as separate calls
// would, and every problem that overlaps nothing earlier shares pass 0.  Empty ranges overlap
// nothing.  Returns the number of passes; wave_of[index] is each problem's pass.
inline int assign_reduce_waves_spans(const std::vector<RhoSpan> &spans, std::vector<int> &wave_of) {
    const size_t n = spans.size();
    wave_of.assign(n, 0);
    if (n == 0) return 1;
    std::vector<RhoSpan> by_lo(spans);
    std::sort(by_lo.begin(), by_lo.end(), [](const RhoSpan &a, const RhoSpan &b) {
        return a.lo != b.lo ? a.lo < b.lo : a.index < b.index;
    });
    auto overlap = [](const RhoSpan &a, const RhoSpan &b) {
        return a.lo < a.hi && b.lo < b.hi && a.lo < b.hi && b.lo < a.hi;
    };
    // prefix maximum of hi over by_lo: span i overlaps an earlier-starting one exactly when it is
    // not empty and pmax[i-1] > its lo (an empty span ends where it starts and reaches no one)
    std::vector<uintptr_t> pmax(n);
    std::vector<size_t> pos(n);
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && by_lo[i].lo < by_lo[i].hi && pmax[i - 1] > by_lo[i].lo) any = true;
        pmax[i] = std::max(i ? pmax[i - 1] : (uintptr_t)0, by_lo[i].hi);
        pos[(size_t)by_lo[i].index] = i;
    }
    if (!any) return 1;
    int nwaves = 1;
    for (size_t idx = 0; idx < n; ++idx) {  // call order: every earlier problem has its pass
        const size_t p = pos[idx];
        const RhoSpan me = by_lo[p];
        int w = 0;
        if (me.lo < me.hi) {
            for (size_t j = p; j-- > 0 && pmax[j] > me.lo;)
                if (overlap(by_lo[j], me) && (size_t)by_lo[j].index < idx)
                    w = std::max(w, wave_of[(size_t)by_lo[j].index] + 1);
            for (size_t j = p + 1; j < n && by_lo[j].lo < me.hi; ++j)
                if (overlap(by_lo[j], me) && (size_t)by_lo[j].index < idx)
                    w = std::max(w, wave_of[(size_t)by_lo[j].index] + 1);
        }
        wave_of[idx] = w;
        nwaves = std::max(nwaves, w + 1);
    }
    return nwaves;
}

}  // namespace cfgmm
