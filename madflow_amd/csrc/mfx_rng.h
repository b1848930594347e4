// Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the same
// (counter, key) gives the same bits on the host and on every lane, so any event of
// any iteration can be regenerated independently (no RNG state in HBM).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfx {

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53-bit uniform in the open interval (0,1)
__host__ __device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Uniforms number 2j and 2j+1 of event ``ev``: counter (ev_lo, ev_hi, j, iteration), key (seed, 0x6d66)
__host__ __device__ __forceinline__ void philox_pair(uint64_t ev, uint32_t j, uint32_t seed, uint32_t iteration,
                                                     double& u0, double& u1) {
  u32x4 c;
  c.x = (uint32_t)ev;
  c.y = (uint32_t)(ev >> 32);
  c.z = j;
  c.w = iteration;
  const u32x4 r = philox4x32_10(c, seed, 0x6d66u);
  u0 = u53(r.x, r.y);
  u1 = u53(r.z, r.w);
}

}  // namespace mfx
