// Counter-based Philox4x32-10 stream shared by the device kernels and the
// host (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11).
//
// One draw = one Philox block with
//   key     = (seed_lo, seed_hi)
//   counter = (draw_lo, draw_hi, stream_lo, stream_hi)
// and the f64 is the top 53 bits of (word1 << 32 | word0) scaled by 2^-53,
// giving a value in [0, 1) with the same resolution as rand's Standard f64.
// stream = shot index (terminal sampling, trajectories, result counts),
// draw   = measurement index inside that stream.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define DQ_HD __host__ __device__ __forceinline__
#else
#define DQ_HD inline
#endif

namespace dq {

DQ_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)M0 * c[0];
    uint64_t p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

DQ_HD double uniform53(uint64_t seed, uint64_t stream, uint64_t draw) {
  uint32_t c[4] = {(uint32_t)draw, (uint32_t)(draw >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t bits = ((uint64_t)c[1] << 32) | c[0];
  return (double)(bits >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace dq
