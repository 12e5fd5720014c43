// Seeded counter-based generator of the synthetic inputs, device side (the
// host side is in nm_setup.cpp): 64-bit mix of seed and index, uniform in
// [0,1) from the top 53 bits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nm {

__device__ inline uint64_t dmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ inline double drng_uniform(uint64_t seed, uint64_t idx) {
  const uint64_t k = dmix64(seed + 0x9e3779b97f4a7c15ULL);
  const uint64_t u = dmix64(k ^ (idx * 0xd1b54a32d192ed03ULL + 0x9e3779b97f4a7c15ULL));
  return (double)(u >> 11) * 0x1.0p-53;
}

}  // namespace nm
