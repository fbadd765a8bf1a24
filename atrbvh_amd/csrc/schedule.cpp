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
