// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Every random draw in
// the library (sample positions, synthetic inputs) is a pure function of
// (counter, key), so runs are reproducible and identical at any launch shape.
#pragma once
#include <stdint.h>

#ifndef GGMS_HD
#if defined(__HIPCC__)
#define GGMS_HD __host__ __device__ __forceinline__
#else
#define GGMS_HD inline
#endif
#endif

namespace ggms {

struct U32x4 {
  uint32_t v[4];
};

GGMS_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// ctr: 4 words, key: 2 words; 10 rounds with the published multipliers and
// Weyl key increments.
GGMS_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c0, hi0, lo0);
    philox_mulhilo(M1, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0;
    k1 += W1;
  }
  U32x4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

// Library-wide key derivation: the 64-bit user seed is mixed with the tag
// 0x67676D73 ("ggms").  Distinct purposes use distinct counter word 2
// ("streams") so that, e.g., sample positions never reuse input draws.
constexpr uint32_t kSeedTag = 0x67676D73u;
enum Stream : uint32_t {
  kStreamValues = 0,       // synthetic input values
  kStreamValues2 = 1,      // second draw per element (mixture choice, masks)
  kStreamSample = 2,       // sample positions for the splitter
  kStreamSubSample = 3,    // second-level sample (strided one-GPU distributed form)
  kStreamKs = 4,           // k-rank generators
  kStreamParams = 5,       // per-problem parameters (mixture means, sizes, ...)
};

GGMS_HD U32x4 philox_draw(uint64_t seed, uint32_t stream, uint64_t idx) {
  return philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u, (uint32_t)seed ^ kSeedTag,
                       (uint32_t)(seed >> 32));
}

// 24-bit uniform in [0, 1) and (0, 1]; both exact in fp32 and fp64.
GGMS_HD float u24_open1(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }
GGMS_HD double u53(uint32_t hi, uint32_t lo) {
  uint64_t m = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
  return (double)m * (1.0 / 9007199254740992.0);
}

}  // namespace ggms
