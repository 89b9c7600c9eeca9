// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11)
// and the draw addressing used by every random draw of the sensor.
//
// A draw is a pure function of (run seed, purpose tag, camera, pixel, 64-bit physical index): the
// counter is {purpose << 8 | camera, pixel, index lo, index hi} and the key is the seed's two words.
// This is the reference's layout (src/core/rng/purpose.h:45-53) so that identical inputs give the
// reference's streams; the generator itself is the published Philox4x32 with 10 rounds, pinned by
// the Random123 known-answer vectors in the tests.
#pragma once

#include "hd.h"

namespace bsx {

struct Word4 {
  uint32_t v[4];
};

// Purpose tags (values fixed by the reference's registry, src/core/rng/purpose.h:14-22).
enum : uint32_t {
  kPurposeSensorMap = 0x0002,
  kPurposeNoise = 0x0010,
  kPurposeNoiseInit = 0x0011,
  kPurposeReadoutPhase = 0x0022,
  // This library's own streams (no counterpart in the reference): the rate-specified shot-noise
  // process of the sensor path. Chosen outside the reference's registry.
  kPurposeShot = 0x0101,
};

BSX_HD BSX_INLINE uint32_t umulhi32(uint32_t a, uint32_t b) {
  return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}

BSX_HD BSX_INLINE Word4 philox4x32_10(Word4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kMulA = 0xD2511F53u, kMulB = 0xCD9E8D57u;
  constexpr uint32_t kBumpA = 0x9E3779B9u, kBumpB = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint32_t ha = umulhi32(kMulA, c.v[0]), la = kMulA * c.v[0];
    const uint32_t hb = umulhi32(kMulB, c.v[2]), lb = kMulB * c.v[2];
    c = Word4{{hb ^ c.v[1] ^ k0, lb, ha ^ c.v[3] ^ k1, la}};
    k0 += kBumpA;
    k1 += kBumpB;
  }
  return c;
}

BSX_HD BSX_INLINE Word4 draw(uint64_t seed, uint32_t purpose, uint32_t camera, uint32_t pixel, uint64_t index) {
  const Word4 ctr{{(purpose << 8) | (camera & 0xFFu), pixel, static_cast<uint32_t>(index),
                   static_cast<uint32_t>(index >> 32)}};
  return philox4x32_10(ctr, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

// A Philox word as a uniform in the open interval (0, 1): 32 bits in fp64, 23 bits in fp32 (so the
// value is exact in a float).
BSX_HD BSX_INLINE double open_uniform64(uint32_t u) { return (static_cast<double>(u) + 0.5) * 0x1p-32; }
BSX_HD BSX_INLINE float open_uniform32(uint32_t u) { return (static_cast<float>(u >> 9) + 0.5f) * 0x1p-23f; }

}  // namespace bsx
