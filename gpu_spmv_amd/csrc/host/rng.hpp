// xoshiro256** (Blackman & Vigna, 2018) seeded through splitmix64.
// Streams: the state of stream s of seed k is four successive splitmix64
// outputs started at k ^ (s * 0x9E3779B97F4A7C15).  Row-parallel generators
// give every row its own stream, so results do not depend on thread count.
#pragma once

#include <cmath>
#include <cstdint>

namespace axspmv {

inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    bool has_spare = false;
    double spare = 0.0;

    Rng(uint64_t seed, uint64_t stream) {
        uint64_t sm = seed ^ (stream * 0x9E3779B97F4A7C15ull);
        for (auto &w : s) w = splitmix64(sm);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // U[0,1) with 53 random bits.
    inline double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    inline double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    // Integer uniform in [0, n) by Lemire's multiply-shift with rejection.
    inline uint64_t below(uint64_t n) {
        uint64_t x = next();
        __uint128_t m = (__uint128_t)x * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            uint64_t t = (0 - n) % n;
            while (l < t) {
                x = next();
                m = (__uint128_t)x * n;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    // N(0,1) by Marsaglia's polar method (pairs; the second is kept).
    inline double normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u, v, q;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            q = u * u + v * v;
        } while (q >= 1.0 || q == 0.0);
        const double m = std::sqrt(-2.0 * std::log(q) / q);
        spare = v * m;
        has_spare = true;
        return u * m;
    }
};

}  // namespace axspmv
