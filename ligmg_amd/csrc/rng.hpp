// Counter-based RNG shared by the host generators and the device setup.
//
// The reference draws its test vectors and right-hand sides from rand()
// seeded with the wall clock (include/ligmg/combblas_extra.hpp:16-22,
// src/main.cpp:26), so no run is reproducible.  Here every random value is a
// pure function of (seed, stream, index): the splitmix64 finaliser (Vigna,
// public domain) applied twice, which lets any thread draw any element
// independently and gives the same bits on the CPU and on the GPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LIGMG_HD __host__ __device__ __forceinline__
#else
#define LIGMG_HD inline
#endif

namespace ligmg_rng {

LIGMG_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

LIGMG_HD uint64_t ctr_hash(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed ^ mix64(stream)) + idx);
}

// Uniform in [0,1) with 53 random bits.
LIGMG_HD double u01(uint64_t seed, uint64_t stream, uint64_t idx) {
  return (double)(ctr_hash(seed, stream, idx) >> 11) * (1.0 / 9007199254740992.0);
}

// Test-vector entries are 32-bit integers k scaled by 2^-32, so that the
// mean of a test vector can be formed exactly with integer sums (the
// reference subtracts the mean, include/ligmg/util.hpp:19-24).
LIGMG_HD uint32_t tv_bits(uint64_t seed, uint64_t level, uint64_t t, uint64_t i) {
  return (uint32_t)(ctr_hash(seed, 0x5456000000000000ull | (level << 16) | t, i) >> 32);
}

// Streams used by the solver's own draws.
static const uint64_t STREAM_EIG = 0x4549470000000000ull;  // eigen-estimate start vector

}  // namespace ligmg_rng
