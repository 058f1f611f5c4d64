// Counter-based random draws for the synthetic inputs and the thermal injection.
//
// SplitMix64's output function (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) applied to a Weyl sequence indexed by (particle, draw): draw j of
// item p under seed s is mix64(s + (16 p + j + 1) * 0x9E3779B97F4A7C15) scaled to [0, 1) by its
// top 53 bits.  Any draw is a pure function of (s, p, j), so a run's numbers do not depend on
// the order or the threads that take them.  Normals come in pairs from Box-Muller (Box and
// Muller 1958) on draws j, j + 1.
#pragma once
#include <cmath>
#include <cstdint>

namespace mprng {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double draw(uint64_t seed, int64_t p, int j) {
    uint64_t z = mix64(seed + ((uint64_t)p * 16u + (uint64_t)j + 1u) * 0x9E3779B97F4A7C15ull);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
}

// Box-Muller pair from draws j, j+1 of item p.
inline void normal2(uint64_t seed, int64_t p, int j, double& z0, double& z1) {
    double u1 = 1.0 - draw(seed, p, j);  // (0,1]
    double u2 = draw(seed, p, j + 1);
    double r = std::sqrt(-2.0 * std::log(u1));
    double a = 2.0 * M_PI * u2;
    z0 = r * std::cos(a);
    z1 = r * std::sin(a);
}

}  // namespace mprng
