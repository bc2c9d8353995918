// ed_rng.hpp -- counter-based N(0,1) draws shared by the device kernels.
//
// SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) hashes (seed, pair index) into two 53-bit
// uniforms; the Box-Muller transform turns them into the pair
// (r cos t, r sin t).  Element i takes component i % 2 of pair i / 2, so a
// vector depends only on (seed, i), not on the launch shape or on which
// kernel draws it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ed {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// element i of the N(0,1) vector of 'seed'
__device__ __forceinline__ double normal_draw(uint64_t seed, uint64_t i) {
  const uint64_t key = splitmix64(seed);
  const uint64_t pr = i >> 1;
  const uint64_t a = splitmix64(key ^ (2 * pr)), b = splitmix64(key ^ (2 * pr + 1));
  const double u1 = ((a >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  const double u2 = (b >> 11) * 0x1.0p-53;        // [0, 1)
  const double r = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  return (i & 1) ? r * sn : r * cs;
}

}  // namespace ed
