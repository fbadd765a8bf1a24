// Seeded synthetic scenes (see include/atrbvh_scenes.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/atrbvh_scenes.h"
#include "internal.hpp"

namespace {

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// PCG32, XSH-RR output (O'Neill 2014).
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed)
    {
        uint64_t s = seed;
        state = splitmix64(s);
        inc = splitmix64(s) | 1ull;
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    double normal()
    {
        const double u1 = ((double)next() + 1.0) / 4294967296.0;
        const double u2 = (double)next() / 4294967296.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint32_t below(uint32_t k) { return (uint32_t)(((uint64_t)next() * k) >> 32); }
};

}  // namespace

extern "C" void atr_gen_soup(uint64_t seed, uint32_t n, float sigma, float* v)
{
    Pcg32 g(seed);
    for (uint32_t i = 0; i < n; ++i) {
        const float c[3] = {g.uniform(), g.uniform(), g.uniform()};
        for (int k = 0; k < 3; ++k) {
            float* p = v + ((size_t)i * 3 + k) * 4;
            for (int a = 0; a < 3; ++a) p[a] = c[a] + (float)((double)sigma * g.normal());
            p[3] = 0.0f;
        }
    }
}

extern "C" void atr_gen_gmm(uint64_t seed, uint32_t n, uint32_t clusters, float sigma_lo,
                            float sigma_hi, float edge_median, float edge_sigma, float* v)
{
    Pcg32 g(seed);
    std::vector<double> centre(3 * (size_t)clusters), sig(clusters);
    const double llo = std::log((double)sigma_lo), lhi = std::log((double)sigma_hi);
    for (uint32_t c = 0; c < clusters; ++c) {
        for (int a = 0; a < 3; ++a) centre[3 * c + a] = g.uniform();
        sig[c] = std::exp(llo + (lhi - llo) * g.uniform());
    }
    const double lmed = std::log((double)edge_median);
    // unit-side equilateral triangle centred at the origin in the xy plane
    const double kPi = 3.141592653589793;
    double P[3][3];
    for (int k = 0; k < 3; ++k) {
        const double ang = 2.0 * kPi * k / 3.0;
        P[k][0] = std::cos(ang) / std::sqrt(3.0);
        P[k][1] = std::sin(ang) / std::sqrt(3.0);
        P[k][2] = 0.0;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = g.below(clusters);
        double cen[3];
        for (int a = 0; a < 3; ++a) cen[a] = centre[3 * c + a] + sig[c] * g.normal();
        const double L = std::exp(lmed + (double)edge_sigma * g.normal());
        double q[4];
        double qn = 0.0;
        for (int a = 0; a < 4; ++a) {
            q[a] = g.normal();
            qn += q[a] * q[a];
        }
        qn = std::sqrt(qn);
        for (int a = 0; a < 4; ++a) q[a] /= qn;
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        for (int k = 0; k < 3; ++k) {
            float* p = v + ((size_t)i * 3 + k) * 4;
            for (int a = 0; a < 3; ++a) {
                const double r = R[a][0] * P[k][0] + R[a][1] * P[k][1] + R[a][2] * P[k][2];
                p[a] = (float)(cen[a] + L * r);
            }
            p[3] = 0.0f;
        }
    }
}

extern "C" void atr_gen_batch_sizes(uint64_t seed, uint32_t meshes, uint32_t lo, uint32_t hi,
                                    uint32_t* offsets)
{
    Pcg32 g(seed);
    const double llo = std::log((double)lo), lhi = std::log((double)hi);
    offsets[0] = 0;
    for (uint32_t m = 0; m < meshes; ++m) {
        double k = std::floor(std::exp(llo + (lhi - llo) * g.uniform()) + 0.5);
        k = std::min(std::max(k, (double)lo), (double)hi);
        offsets[m + 1] = offsets[m] + (uint32_t)k;
    }
}

extern "C" void atr_gen_soup_batch(uint64_t seed_base, uint32_t meshes, const uint32_t* offsets,
                                   float sigma, float* v)
{
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (uint32_t m = next++; m < meshes; m = next++)
            atr_gen_soup(seed_base + m, offsets[m + 1] - offsets[m], sigma,
                         v + (size_t)offsets[m] * 12);
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < hw; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

extern "C" int atr_trbvh_schedule(int treelet, int warp, int32_t* out, int cap)
{
    std::vector<int32_t> s;
    const int rounds = atr::trbvh_schedule(treelet, warp, s);
    if (rounds < 0 || (int)s.size() > cap) return -1;
    for (size_t i = 0; i < s.size(); ++i) out[i] = s[i];
    return rounds;
}
