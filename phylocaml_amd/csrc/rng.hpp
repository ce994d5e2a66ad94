// Counter-based random numbers for the synthetic workloads (host only).
// SplitMix64 finaliser (Steele, Lea & Flood 2014) over (seed, stream, counter):
// every draw is a pure function of its key, so results do not depend on how
// the work is split across host threads.
#pragma once
#include <cmath>
#include <cstdint>

namespace phy {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t state;
  Rng(uint64_t seed, uint64_t stream) : state(mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull))) {}
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0,1)
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // uniform in (0,1]
  double uniform_pos() { return ((double)(next() >> 11) + 1.0) * 0x1.0p-53; }
  double exponential(double mean) { return -mean * std::log(uniform_pos()); }
  double normal() {
    // Box-Muller
    double u1 = uniform_pos(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Marsaglia & Tsang (2000) gamma(shape, 1)
  double gamma(double shape) {
    if (shape < 1.0) {
      double g = gamma(shape + 1.0);
      return g * std::pow(uniform_pos(), 1.0 / shape);
    }
    const double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x = normal(), v = 1.0 + c * x;
      if (v <= 0.0) continue;
      v = v * v * v;
      double u = uniform_pos();
      if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
  // Knuth's product method; fine for small lambda
  uint32_t poisson(double lambda) {
    const double l = std::exp(-lambda);
    uint32_t k = 0;
    double p = 1.0;
    for (;;) {
      p *= uniform_pos();
      if (p <= l) return k;
      ++k;
    }
  }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

}  // namespace phy
