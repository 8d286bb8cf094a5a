// pairmap.hpp -- the collider-pair keyed persistence map
// (20_broadphase.metal:547-608): open addressing with linear probing; key A
// (= colliderA + 1) is claimed by CAS, key B and the value are written by the
// winner.  Built from one frame's manifolds, queried by the next frame's
// narrow phase and by the debug lookup.
#pragma once
#include "state.hpp"

namespace avbd {

__device__ __forceinline__ uint32_t pair_map_hash(uint32_t a, uint32_t b, uint32_t cap) {
  uint32_t h = (a * 73856093u) ^ (b * 19349663u);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h & (cap - 1u);
}

// index of (a, b) in the previous frame's manifolds, or -1
__device__ __forceinline__ int pair_map_find(const DeviceState& S, uint32_t cap, uint32_t a, uint32_t b) {
  uint32_t h = pair_map_hash(a, b, cap);
  for (uint32_t probe = 0; probe < cap; probe++) {
    uint32_t ka = S.mapKeyA[h];
    if (ka == 0) return -1;
    if (ka == a + 1 && S.mapKeyB[h] == b) return (int)S.mapVal[h];
    h = (h + 1) & (cap - 1u);
  }
  return -1;
}

}  // namespace avbd
