// galic_rng.h -- the counter-based generator of the device draws.
//
// Philox4x32-10 (Salmon et al., SC'11) with the counter layout of
// galic_amd/rng.py: (id lo, id hi, attempt, 2 * block + half) under the key
// (seed, stream).  Every draw is a pure function of its counter, so a sample's
// numbers do not depend on which thread or launch makes it, and the host module
// gives the same values.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace galic {

__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                             uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
  uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
  uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
  c0 = n0;
  c1 = lo1;
  c2 = n2;
  c3 = lo0;
}

__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for(int r = 0; r < 10; r++) {
    philox_round(c[0], c[1], c[2], c[3], k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// A uniform double in (0, 1) from two 32-bit words (53 random bits, as rng.py).
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) / 9007199254740992.0;
}

// Four uniforms for (id, attempt, block), as rng.uniforms
__device__ __forceinline__ void uniform4(uint64_t id, uint32_t attempt, uint32_t block, uint32_t seed, uint32_t stream,
                                         double u[4]) {
  for(int half = 0; half < 2; half++) {
    uint32_t c[4] = {(uint32_t)id, (uint32_t)(id >> 32), attempt, 2 * block + half};
    philox4x32(c, seed, stream);
    u[2 * half] = u01(c[0], c[1]);
    u[2 * half + 1] = u01(c[2], c[3]);
  }
}

// Four standard normals for (id, attempt, block), as rng.gaussians (Box-Muller)
__device__ __forceinline__ void gauss4(uint64_t id, uint32_t attempt, uint32_t block, uint32_t seed, uint32_t stream,
                                       double g[4]) {
  double u[4];
  uniform4(id, attempt, block, seed, stream, u);
  double r1 = sqrt(-2.0 * log(u[0])), r2 = sqrt(-2.0 * log(u[2]));
  g[0] = r1 * cos(2 * M_PI * u[1]);
  g[1] = r1 * sin(2 * M_PI * u[1]);
  g[2] = r2 * cos(2 * M_PI * u[3]);
  g[3] = r2 * sin(2 * M_PI * u[3]);
}

}  // namespace galic
