// Philox4x32-10 and Philox2x32-10 counter-based generators (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Ten rounds of the
// Philox S-box with the published multipliers and Weyl key increments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecsprt {

struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c.v[0], hi0, lo0);
    philox_mulhilo(M1, c.v[2], hi1, lo1);
    u32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

struct u32x2 {
  uint32_t v[2];
};

__host__ __device__ inline u32x2 philox2x32_10(u32x2 c, uint32_t k) {
  const uint32_t M = 0xD256D193u, W = 0x9E3779B9u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi, lo;
    philox_mulhilo(M, c.v[0], hi, lo);
    u32x2 n;
    n.v[0] = hi ^ k ^ c.v[1];
    n.v[1] = lo;
    c = n;
    k += W;
  }
  return c;
}

// Uniform float in (0, 1): top 24 bits, centred in their bin.
__host__ __device__ inline float u01(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }

}  // namespace ecsprt
