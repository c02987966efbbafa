// Counter-based random numbers for the frame kernels: every sample is a pure
// function of (seed, frame, ray index, dimension), so results do not depend
// on launch shape or scheduling.  Mixing function: the "lowbias32" integer
// hash published by C. Wellons (hash-prospector), applied in a chain.
#ifndef RNG_H
#define RNG_H

#include <stdint.h>

__host__ __device__ __forceinline__ uint32_t rng_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__host__ __device__ __forceinline__ uint32_t rng_key(uint32_t seed, uint32_t frame,
                                                     uint32_t idx) {
  return rng_hash(seed ^ rng_hash(frame ^ rng_hash(idx + 0x632be5abu)));
}

// Uniform in [0, 1) with 24 random bits.
__host__ __device__ __forceinline__ float rng_u01(uint32_t key, uint32_t dim) {
  return (float)(rng_hash(key + 0x9e3779b9u * (dim + 1u)) >> 8) * (1.0f / 16777216.0f);
}

#endif
