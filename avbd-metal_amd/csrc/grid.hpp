// grid.hpp -- uniform-grid cell coordinates and the avalanche cell hash
// (20_broadphase.metal:12-32), shared by the rigid broad phase and the cloth
// element grid.  Integer results must be reproducible bit for bit, so
// floating-point contraction is off for every file that includes this.
#pragma once
#include "state.hpp"

#pragma clang fp contract(off)

namespace avbd {

__device__ __forceinline__ void cell_coord(V3 p, float cellSize, int& cx, int& cy, int& cz) {
  cx = (int)floorf(p.x / cellSize);
  cy = (int)floorf(p.y / cellSize);
  cz = (int)floorf(p.z / cellSize);
}
// Multiply-xor of three primes, salted by collision group, then an
// avalanche finish (xorshift-multiply) so lattice-aligned cells spread.
__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t group, uint32_t hashSize) {
  uint32_t h = ((uint32_t)cx * 92837111u) ^ ((uint32_t)cy * 689287499u) ^ ((uint32_t)cz * 283923481u);
  h ^= group * 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h & (hashSize - 1u);
}

// sorted exclusion list (body pairs, lo < hi): binary search
__device__ __forceinline__ bool pair_excluded(const DeviceState& S, uint32_t count, uint32_t a, uint32_t b) {
  uint32_t lo = 0, hi = count;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    U2 e = S.exclusions[mid];
    if (e.x < a || (e.x == a && e.y < b)) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= count) return false;
  U2 e = S.exclusions[lo];
  return e.x == a && e.y == b;
}

}  // namespace avbd
