// Counter-based RNG shared by the host setup library and the GPU library: SplitMix64's
// output function applied to a Weyl sequence, addressed by (seed, stream, counter), so the
// same draw can be produced by any consumer (generators, Lanczos start vectors, tests).
#pragma once

#include <stdint.h>

namespace chrng {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ULL;

// Draw k of stream `stream` under `seed`.
inline uint64_t u64(uint64_t seed, uint64_t stream, uint64_t k) {
  const uint64_t base = mix64(seed ^ (stream * 0xd1b54a32d192ed03ULL));
  return mix64(base + (k + 1) * kGamma);
}

// Uniform in [0, 1) from the top 53 bits.
inline double u01(uint64_t seed, uint64_t stream, uint64_t k) {
  return (double)(u64(seed, stream, k) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace chrng
