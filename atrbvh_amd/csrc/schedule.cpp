// Host-generated lane schedule for the TRBVH subset DP
// (TRBVHScheduler::GenerateSchedule, BVHRT-Core/src/TRBVHScheduler.cpp:29-126).
//
// The reference records, by recursive partitioning from the full set, which
// supersets each subset of size 2..n-2 is partitioned from.  Every subset is
// reachable and every set of size >= 3 gets expanded, so a subset's
// dependencies are exactly its strict supersets.  Subsets are taken by size
// from n-2 down to 2 and by ascending value, each placed in the first round
// after all its dependencies that has a free lane; rounds are padded with 0
// and reversed so smaller subsets run first.
#include <algorithm>
#include <cstdint>
#include <vector>

#include <utility>

#include "internal.hpp"

namespace atr {

int trbvh_schedule(int treelet, int warp, std::vector<int32_t>& out)
{
    if (treelet < 3 || treelet > 16 || warp <= 0) return -1;
    const int nsub = 1 << treelet;
    std::vector<int> round_of(nsub, -1);
    std::vector<std::vector<int32_t>> rounds;
    for (int size = treelet - 2; size >= 2; --size) {
        for (int s = 1; s < nsub; ++s) {
            if (__builtin_popcount((unsigned)s) != size) continue;
            // earliest round after every strict superset already placed
            int minimum = 0;
            const int free_bits = (nsub - 1) & ~s;
            for (int extra = free_bits; extra != 0; extra = (extra - 1) & free_bits) {
                const int sup = s | extra;
                if (round_of[sup] >= 0) minimum = std::max(minimum, round_of[sup] + 1);
            }
            int r = minimum;
            while (true) {
                if ((int)rounds.size() <= r) rounds.emplace_back();
                if ((int)rounds[r].size() < warp) break;
                ++r;
            }
            rounds[r].push_back(s);
            round_of[s] = r;
        }
    }
    out.assign((size_t)rounds.size() * warp, 0);
    for (size_t r = 0; r < rounds.size(); ++r) {
        const size_t dst = rounds.size() - 1 - r;
        std::copy(rounds[r].begin(), rounds[r].end(), out.begin() + dst * warp);
    }
    return (int)rounds.size();
}

}  // namespace atr

namespace atr {

void host_spans(uint32_t n, const int32_t* left, const int32_t* right, int2* span)
{
    const uint32_t last = n - 1;
    for (uint32_t i = last; i < 2 * n - 1; ++i) span[i] = make_int2((int)(i - last), (int)(i - last));
    // iterative post-order from the root
    std::vector<std::pair<int32_t, int32_t>> st;
    st.emplace_back(0, 0);
    while (!st.empty()) {
        auto& f = st.back();
        const int32_t node = f.first;
        if ((uint32_t)node >= last) {
            st.pop_back();
            continue;
        }
        if (f.second == 0) {
            f.second = 1;
            st.emplace_back(right[node], 0);
            st.emplace_back(left[node], 0);
            continue;
        }
        st.pop_back();
        const int2 a = span[left[node]];
        const int2 b = span[right[node]];
        span[node] = make_int2(std::min(a.x, b.x), std::max(a.y, b.y));
    }
}

}  // namespace atr
