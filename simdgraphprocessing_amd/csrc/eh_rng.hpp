// eh_rng.hpp -- counter-based random draws for the seeded graph generators.
//
//   mix64(z)        SplitMix64 output finaliser
//   stream_key(s,t) mix64(s + mix64(t))
//   draw(k, c)      mix64(k + (c + 1) * 0x9E3779B97F4A7C15)
//
// Every draw is a pure function of (seed, stream, counter), so any thread can
// generate any edge independently and the host and the device agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define EH_GOLDEN 0x9E3779B97F4A7C15ull

__host__ __device__ inline uint64_t eh_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t eh_stream_key(uint64_t seed, uint64_t stream) {
  return eh_mix64(seed + eh_mix64(stream));
}

__host__ __device__ inline uint64_t eh_draw(uint64_t key, uint64_t ctr) {
  return eh_mix64(key + (ctr + 1) * EH_GOLDEN);
}

// map a uniform 32-bit value onto [0, n)
__host__ __device__ inline uint32_t eh_scale32(uint32_t r, uint32_t n) {
  return (uint32_t)(((uint64_t)r * n) >> 32);
}

// Batch member g: n uniform in [64, 8192], mean degree uniform in [4, 64),
// m draws = n * degree / 2 (integer fixed point, no floating point).
__host__ __device__ inline void eh_batch_params(uint64_t seed, uint32_t g, uint32_t *n_out,
                                                uint64_t *m_out) {
  uint64_t r = eh_draw(eh_stream_key(seed, 0), g);
  uint32_t n = 64 + (uint32_t)(((uint64_t)(uint32_t)r * 8129ull) >> 32);
  uint64_t deg_q32 = 4ull * 4294967296ull + 60ull * (uint64_t)(uint32_t)(r >> 32);
  *n_out = n;
  *m_out = ((uint64_t)n * deg_q32) >> 33;
}
