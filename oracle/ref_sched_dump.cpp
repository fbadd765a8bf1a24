// Driver for the reference TRBVH scheduler compiled from its own source:
// prints "rounds warp" then every schedule entry, one round per line.
// Usage: trbvh_sched_dump <treelet> <warp>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "TRBVHScheduler.h"

int main(int argc, char** argv)
{
    int treelet = argc > 1 ? std::atoi(argv[1]) : 7;
    int warp = argc > 2 ? std::atoi(argv[2]) : 32;
    BVHRT::TRBVHScheduler scheduler;
    std::vector<std::vector<int>> schedule;
    scheduler.GenerateSchedule(treelet, warp, schedule);
    std::printf("%d %d\n", (int)schedule.size(), warp);
    for (const auto& round : schedule) {
        for (size_t i = 0; i < round.size(); ++i) std::printf(i ? " %d" : "%d", round[i]);
        std::printf("\n");
    }
    return 0;
}
