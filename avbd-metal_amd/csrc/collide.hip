// collide.hip -- spatial hash, deterministic rigid pair stream, CAS
// persistence map and narrow phase for gfx950.
//
// Integer outputs of this file (hash cells, cell lists, pair sets, map
// contents) must be bit-reproducible, so floating-point contraction is off
// here: every product/sum rounds exactly as written.
#include "kernels.hpp"
#include "grid.hpp"

#pragma clang fp contract(off)

namespace avbd {

// ---------------------------------------------------------------------------
// spatial hash (20_broadphase.metal:12-32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 collider_center(const DeviceState& S, uint32_t c) {
  uint32_t b = S.colliderOwner[c];
  return xyz(S.posLin[b]) + q_rotate(S.posAng[b], xyz(S.colliderLocalPos[c]));
}
__device__ __forceinline__ bool groups_compatible(uint32_t a, uint32_t b) {
  return a == 0u || b == 0u || a == b;
}
__device__ __forceinline__ bool domains_compatible(uint32_t ga, uint32_t gb, uint32_t sa, uint32_t sb) {
  if (!groups_compatible(ga, gb)) return false;
  if (ga == 0u && gb != 0u) return sb != 0u;
  if (gb == 0u && ga != 0u) return sa != 0u;
  return true;
}

// Count colliders per cell and remember each collider's (cell, slot).
__global__ void __launch_bounds__(256) k_bp_count(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numHashed) return;
  uint32_t c = S.hashedIdx[gid];
  V3 p = collider_center(S, c);
  int cx, cy, cz;
  cell_coord(p, P.cellSize, cx, cy, cz);
  uint32_t h = cell_hash(cx, cy, cz, S.colliderGroup[c], P.gridHashSize);
  uint32_t slot = atomicAdd(&S.cellCount[h], 1u);
  S.bodyCellSlot[gid] = U2{h, slot};
  float r = S.colliderShape[c].w;
  S.hashCenter[gid] = f4(p, fabsf(r));
  if (r >= 0.0f) S.cellRigid[h] = 1u;
}

__global__ void __launch_bounds__(256) k_bp_scatter(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numHashed) return;
  U2 cs = S.bodyCellSlot[gid];
  S.cellBodies[S.cellStart[cs.x] + cs.y] = S.hashedIdx[gid];
}

// Sort each bucket by collider id so pair enumeration order is independent
// of the atomic scatter schedule.  Buckets of up to 32 (hash aliasing puts
// two or three lattice cells in one bucket) are sorted in registers by an
// odd-even transposition network; longer ones in place.
__global__ void __launch_bounds__(256) k_bp_sort_cells(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.gridHashSize) return;
  uint32_t n = S.cellCount[gid];
  if (n < 2) return;
  uint32_t s = S.cellStart[gid];
  if (n <= 32) {
    uint32_t v[32];
#pragma unroll
    for (int i = 0; i < 32; i++) v[i] = (uint32_t)i < n ? S.cellBodies[s + i] : 0xFFFFFFFFu;
#pragma unroll
    for (int pass = 0; pass < 32; pass++) {
#pragma unroll
      for (int i = pass & 1; i + 1 < 32; i += 2) {
        uint32_t a = v[i], b = v[i + 1];
        v[i] = a < b ? a : b;
        v[i + 1] = a < b ? b : a;
      }
    }
#pragma unroll
    for (int i = 0; i < 32; i++)
      if ((uint32_t)i < n) S.cellBodies[s + i] = v[i];
    return;
  }
  // longer buckets (dense particle packings under the 0.05 m radius floor
  // of the cell size): one workgroup each, k_bp_sort_large
  S.largeCells[atomicAdd(&S.counters[CTR_LARGE_CELLS], 1u)] = gid;
}

// Buckets of more than 32 colliders, one workgroup per bucket: bitonic sort
// in LDS (up to kSortLds entries; a larger bucket is insertion-sorted by one
// thread).
constexpr uint32_t kSortLds = 4096;
__global__ void __launch_bounds__(256) k_bp_sort_large(DeviceState S, StepParams P) {
  __shared__ uint32_t v[kSortLds];
  const uint32_t count = S.counters[CTR_LARGE_CELLS];
  const uint32_t tid = threadIdx.x;
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t h = S.largeCells[q];
    const uint32_t n = S.cellCount[h], s0 = S.cellStart[h];
    if (n > kSortLds) {
      if (tid == 0) {
        for (uint32_t i = 1; i < n; i++) {
          uint32_t x = S.cellBodies[s0 + i];
          uint32_t j = i;
          while (j > 0 && S.cellBodies[s0 + j - 1] > x) {
            S.cellBodies[s0 + j] = S.cellBodies[s0 + j - 1];
            j--;
          }
          S.cellBodies[s0 + j] = x;
        }
      }
      __syncthreads();
      continue;
    }
    uint32_t m = 64;
    while (m < n) m <<= 1;
    for (uint32_t i = tid; i < m; i += blockDim.x) v[i] = i < n ? S.cellBodies[s0 + i] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = tid; i < m; i += blockDim.x) {
          const uint32_t l = i ^ j;
          if (l > i) {
            const uint32_t x = v[i], y = v[l];
            if ((x > y) == ((i & k) == 0u)) {
              v[i] = y;
              v[l] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = tid; i < n; i += blockDim.x) S.cellBodies[s0 + i] = v[i];
    __syncthreads();
  }
}

// Candidate search grid: cells of 2 x the largest hashed radius, keyed by the
// same hash and collision group; holds hashed-collider indices, unordered.
__global__ void __launch_bounds__(256) k_bp_fine_count(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numHashed) return;
  int fx, fy, fz;
  cell_coord(xyz(S.hashCenter[gid]), P.fineCellSize, fx, fy, fz);
  uint32_t h = cell_hash(fx, fy, fz, S.colliderGroup[S.hashedIdx[gid]], P.gridHashSize);
  S.fineSlot[gid] = U2{h, atomicAdd(&S.fineCount[h], 1u)};
}

__global__ void __launch_bounds__(256) k_bp_fine_scatter(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numHashed) return;
  U2 fs = S.fineSlot[gid];
  S.fineBodies[S.fineStart[fs.x] + fs.y] = gid;
}

// cell-ordered copies of the per-collider center / metadata
__global__ void __launch_bounds__(256) k_bp_gather(DeviceState S, StepParams P) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.numHashed) return;
  uint32_t g = S.colliderToHashed[S.cellBodies[k]];
  S.cellPos[k] = S.hashCenter[g];
  S.cellMeta[k] = S.hashMeta[g];
}

// Candidate filters shared by every producer kind (20_broadphase.metal:287-424).
__device__ __forceinline__ bool pair_accept(const DeviceState& S, const StepParams& P, uint32_t ownerA, uint32_t grpA,
                                            uint32_t shA, bool aDyn, V3 pa, float ra, uint32_t ownerB,
                                            uint32_t grpB, uint32_t shB, bool bDyn, V3 pb, float rb) {
  if (!domains_compatible(grpA, grpB, shA, shB)) return false;
  if (ownerA == ownerB) return false;
  if (!aDyn && !bDyn) return false;
  V3 d = pa - pb;
  float r = ra + rb;
  if (dot(d, d) > r * r) return false;
  uint32_t lo = ownerA < ownerB ? ownerA : ownerB, hi = ownerA < ownerB ? ownerB : ownerA;
  return !pair_excluded(S, P.numExclusions, lo, hi);
}

// LDS writes of one lane visible to the other lanes of its wave
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kFineCap = 128;  // accepted candidates one wave sorts in LDS

// One wave per producer (a hashed collider scanning its 27 neighbour cells
// and the global list, or a global collider scanning the later globals).
// Lanes split the producer's linearised candidate list; a ballot ranks the
// accepted candidates so the emission order equals the serial order --
// cells in (dz, dy, dx) order, each bucket in collider order, then globals.
// Pass 1 (EMIT = false) writes the per-producer count, pass 2 emits at the
// scanned offset.
// With the candidate search grid (P.fineCellSize > 0: hash cells much larger
// than the colliders), a hashed producer instead scans its 27 search cells,
// applies the same filters, keys every accepted candidate by (position of
// its hash bucket in the 27-cell visit order, collider id) and emits the
// keys sorted -- the same set in the same order as the bucket walk, since an
// accepted partner lies within the radius sum, hence in an adjacent hash
// cell.  More than kFineCap accepted candidates fall back to the walk.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_bp_pairs(DeviceState S, StepParams P) {
  __shared__ uint64_t fineKeys[4][2][kFineCap];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t producer = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (producer >= P.numHashed + P.numGlobals) return;  // whole wave
  const uint64_t ltMask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  const uint32_t base = EMIT ? S.pairStart[producer] : 0u;
  uint32_t count = 0;
  auto emitRanked = [&](bool accept, uint32_t a, uint32_t b) {
    uint64_t mask = __ballot(accept);
    if (EMIT && accept) {
      uint32_t slot = base + count + (uint32_t)__builtin_popcountll(mask & ltMask);
      if (slot < P.maxPairs) S.pairs[slot] = U2{a, b};
    }
    count += (uint32_t)__builtin_popcountll(mask);
  };
  if (producer < P.numHashed) {
    const uint32_t a = S.hashedIdx[producer];
    const U4 ma = S.hashMeta[producer];
    const F4 ca = S.hashCenter[producer];
    const uint32_t ownerA = ma.x, grpA = ma.y, shA = ma.z & 1u, ga = ma.w;
    const bool aDyn = (ma.z & 2u) != 0;
    const V3 pa = xyz(ca);
    const float ra = ca.w;
    if (!(ga != 0 && P.numHashedRigid == 0)) {
      int cx, cy, cz;
      cell_coord(pa, P.cellSize, cx, cy, cz);
      // lanes 0..26 own one neighbour cell each
      uint32_t h = 0, cs = 0, cc = 0;
      if (lane < 27) {
        int q = (int)lane;
        h = cell_hash(cx + q % 3 - 1, cy + (q / 3) % 3 - 1, cz + q / 9 - 1, grpA, P.gridHashSize);
      }
      bool dup = false;
      for (uint32_t j = 0; j < 27; j++) {
        uint32_t hj = __shfl(h, (int)j, 64);
        dup |= (j < lane) && (hj == h);
      }
      bool walk = true;
      if (P.fineCellSize > 0.0f) {
        uint64_t* keys = fineKeys[threadIdx.x >> 6][0];
        uint64_t* sorted = fineKeys[threadIdx.x >> 6][1];
        int fx, fy, fz;
        cell_coord(pa, P.fineCellSize, fx, fy, fz);
        uint32_t hf = 0, fs = 0, fc = 0;
        if (lane < 27) {
          int q = (int)lane;
          hf = cell_hash(fx + q % 3 - 1, fy + (q / 3) % 3 - 1, fz + q / 9 - 1, grpA, P.gridHashSize);
        }
        bool fdup = false;
        for (uint32_t j = 0; j < 27; j++) {
          uint32_t hj = __shfl(hf, (int)j, 64);
          fdup |= (j < lane) && (hj == hf);
        }
        if (lane < 27 && !fdup) {
          fs = S.fineStart[hf];
          fc = S.fineCount[hf];
        }
        uint32_t fincl = fc;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
          uint32_t y = __shfl_up(fincl, d, 64);
          if (lane >= (uint32_t)d) fincl += y;
        }
        const uint32_t fexcl = fincl - fc;
        const uint32_t ftotal = __shfl(fincl, 26, 64);
        uint32_t nAcc = 0;
        for (uint32_t chunk = 0; chunk < ftotal; chunk += 64) {
          const uint32_t j = chunk + lane;
          uint32_t k = 0;
          bool valid = false;
          for (uint32_t t = 0; t < 27; t++) {
            uint32_t et = __shfl(fexcl, (int)t, 64), ct = __shfl(fc, (int)t, 64), st = __shfl(fs, (int)t, 64);
            if (j >= et && j < et + ct) { k = st + (j - et); valid = true; }
          }
          uint32_t b = 0, hb = 0xFFFFFFFFu;
          U4 mb = U4{0, 0, 0, 0};
          F4 cb = F4{0, 0, 0, 0};
          bool cand = false;
          if (valid) {
            const uint32_t g = S.fineBodies[k];
            b = S.hashedIdx[g];
            cand = b > a;
            if (cand) {
              mb = S.hashMeta[g];
              cb = S.hashCenter[g];
              int bx, by, bz;
              cell_coord(xyz(cb), P.cellSize, bx, by, bz);
              hb = cell_hash(bx, by, bz, S.colliderGroup[b], P.gridHashSize);
            }
          }
          uint32_t pos = 27;  // first position of the partner's bucket in the visit order
          for (uint32_t q = 0; q < 27; q++) {
            uint32_t hq = __shfl(h, (int)q, 64);
            if (pos == 27 && hq == hb) pos = q;
          }
          bool accept = false;
          if (cand && pos < 27) {
            accept = !(ga != 0 && S.cellRigid[hb] == 0) && !(ga != 0 && mb.w != 0) &&
                     pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, mb.x, mb.y, mb.z & 1u, (mb.z & 2u) != 0,
                                 xyz(cb), cb.w);
          }
          const uint64_t mask = __ballot(accept);
          if (accept) {
            const uint32_t idx = nAcc + (uint32_t)__builtin_popcountll(mask & ltMask);
            if (idx < kFineCap) keys[idx] = ((uint64_t)pos << 32) | b;
          }
          nAcc += (uint32_t)__builtin_popcountll(mask);
        }
        if (nAcc <= kFineCap) {
          walk = false;
          wave_lds_sync();
          for (uint32_t i = lane; i < nAcc; i += 64) {
            const uint64_t key = keys[i];
            uint32_t rank = 0;
            for (uint32_t j = 0; j < nAcc; j++) rank += keys[j] < key ? 1u : 0u;
            sorted[rank] = key;
          }
          wave_lds_sync();
          for (uint32_t i0 = 0; i0 < nAcc; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool ok = i < nAcc;
            emitRanked(ok, a, ok ? (uint32_t)(sorted[i] & 0xFFFFFFFFull) : 0u);
          }
          wave_lds_sync();  // the next producer of this wave rewrites the keys
        }
      }
      if (walk && lane < 27 && !dup && !(ga != 0 && S.cellRigid[h] == 0)) {
        cs = S.cellStart[h];
        cc = S.cellCount[h];
      }
      uint32_t incl = cc;
#pragma unroll
      for (int d = 1; d < 32; d <<= 1) {
        uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += y;
      }
      const uint32_t excl = incl - cc;
      const uint32_t total = walk ? __shfl(incl, 26, 64) : 0u;
      for (uint32_t chunk = 0; chunk < total; chunk += 64) {
        uint32_t j = chunk + lane;
        uint32_t k = 0;
        bool valid = false;
        for (uint32_t t = 0; t < 27; t++) {
          uint32_t et = __shfl(excl, (int)t, 64), ct = __shfl(cc, (int)t, 64), st = __shfl(cs, (int)t, 64);
          if (j >= et && j < et + ct) { k = st + (j - et); valid = true; }
        }
        bool accept = false;
        uint32_t b = 0;
        if (valid) {
          b = S.cellBodies[k];
          if (b > a) {
            const U4 mb = S.cellMeta[k];
            const F4 cb = S.cellPos[k];
            accept = !(ga != 0 && mb.w != 0) &&
                     pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, mb.x, mb.y, mb.z & 1u, (mb.z & 2u) != 0,
                                 xyz(cb), cb.w);
          }
        }
        emitRanked(accept, a, b);
      }
    }
    for (uint32_t g0 = 0; g0 < P.numGlobals; g0 += 64) {
      uint32_t g = g0 + lane;
      bool accept = false;
      uint32_t b = 0;
      if (g < P.numGlobals) {
        b = S.globalIdx[g];
        uint32_t ownerB = S.colliderOwner[b];
        accept = !(ga != 0 && S.clothGroup[ownerB] == ga) &&
                 pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, ownerB, S.colliderGroup[b],
                             S.colliderShared[b], S.posLin[ownerB].w > 0.0f, collider_center(S, b),
                             fabsf(S.colliderShape[b].w));
      }
      emitRanked(accept, a < b ? a : b, a < b ? b : a);
    }
  } else {
    const uint32_t gp = producer - P.numHashed;
    const uint32_t a = S.globalIdx[gp];
    const uint32_t ownerA = S.colliderOwner[a];
    const uint32_t grpA = S.colliderGroup[a], shA = S.colliderShared[a];
    const V3 pa = collider_center(S, a);
    const float ra = fabsf(S.colliderShape[a].w);
    const bool aDyn = S.posLin[ownerA].w > 0.0f;
    const uint32_t ga = S.clothGroup[ownerA];
    for (uint32_t g0 = gp + 1; g0 < P.numGlobals; g0 += 64) {
      uint32_t g = g0 + lane;
      bool accept = false;
      uint32_t b = 0;
      if (g < P.numGlobals) {
        b = S.globalIdx[g];
        uint32_t ownerB = S.colliderOwner[b];
        accept = !(ga != 0 && S.clothGroup[ownerB] == ga) &&
                 pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, ownerB, S.colliderGroup[b],
                             S.colliderShared[b], S.posLin[ownerB].w > 0.0f, collider_center(S, b),
                             fabsf(S.colliderShape[b].w));
      }
      emitRanked(accept, a < b ? a : b, a < b ? b : a);
    }
  }
  if (!EMIT && lane == 0) S.pairCount[producer] = count;
}

__global__ void k_bp_finalize(DeviceState S, StepParams P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t producers = P.numHashed + P.numGlobals;
  uint32_t n = producers == 0 ? 0 : S.pairStart[producers - 1] + S.pairCount[producers - 1];
  S.counters[CTR_PAIR_CANDIDATES] = n;
  S.counters[CTR_PAIRS] = n < P.maxPairs ? n : P.maxPairs;
}

// ---------------------------------------------------------------------------
// collider-pair keyed persistence map (20_broadphase.metal:547-608):
// open addressing, linear probing; key A (= colliderA + 1) claimed by CAS,
// key B and value written by the winner.  Built from this frame's
// manifolds, queried by next frame's narrow phase.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pair_map_hash(uint32_t a, uint32_t b, uint32_t cap) {
  uint32_t h = (a * 73856093u) ^ (b * 19349663u);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h & (cap - 1u);
}

__global__ void __launch_bounds__(256) k_pm_insert(DeviceState S, StepParams P) {
  uint32_t n = S.counters[CTR_PAIRS];
  for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += gridDim.x * blockDim.x) {
    U4 header = S.manifolds[gid].header;
    if (header.z == 0) continue;
    uint32_t a = S.manifolds[gid].colliderPair.x;
    uint32_t b = S.manifolds[gid].colliderPair.y;
    uint32_t h = pair_map_hash(a, b, P.mapCapacity);
    for (uint32_t probe = 0; probe < P.mapCapacity; probe++) {
      if (atomicCAS(&S.mapKeyA[h], 0u, a + 1u) == 0u) {
        S.mapKeyB[h] = b;
        S.mapVal[h] = gid;
        break;
      }
      h = (h + 1) & (P.mapCapacity - 1u);
    }
  }
}

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

__global__ void k_map_find(DeviceState S, StepParams P, const U2* q, int32_t* out, uint32_t n) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  out[gid] = pair_map_find(S, P.mapCapacity, q[gid].x, q[gid].y);
}

// ---------------------------------------------------------------------------
// narrow phase (30_narrowphase.metal:337-1322): sphere-sphere, sphere-box
// and box-box SAT with reference-face clipping.  Torus / capsule / convex
// hull colliders are rejected at create time.
// ---------------------------------------------------------------------------
constexpr int kMaxPoly = 16;
constexpr float kSatEps = 1.0e-6f;
constexpr float kPlaneEps = 1.0e-5f;
constexpr float kMergeDistSq = 1.0e-6f;

struct NPBox {
  V3 c, h, ax[3];
};
struct NPBest {
  int type, ia, ib;
  float sep;
  V3 n;
  bool valid;
};
struct NPContact {
  V3 xA, xB;
  uint32_t feature;
};

__device__ __forceinline__ float np_half(const NPBox& b, int i) { return comp(b.h, i); }

__device__ __forceinline__ V3 np_support(const NPBox& b, V3 d) {
  float sx = dot(d, b.ax[0]) >= 0.0f ? 1.0f : -1.0f;
  float sy = dot(d, b.ax[1]) >= 0.0f ? 1.0f : -1.0f;
  float sz = dot(d, b.ax[2]) >= 0.0f ? 1.0f : -1.0f;
  return b.c + b.ax[0] * (b.h.x * sx) + b.ax[1] * (b.h.y * sy) + b.ax[2] * (b.h.z * sz);
}

__device__ __forceinline__ float np_spec_cap(float ra, float rb, float margin) {
  float r = fminf(fabsf(ra), fabsf(rb));
  return fminf(0.25f, fmaxf(4.0f * margin, 3.0f * r));
}
__device__ __forceinline__ float np_detect_margin(V3 n, V3 relVel, const StepParams& P, float ra, float rb) {
  float approach = fmaxf(0.0f, -dot(relVel, n));
  float allowed = fminf(approach * P.dt, np_spec_cap(ra, rb, P.collisionMargin));
  return P.collisionMargin + allowed;
}
__device__ __forceinline__ V3 np_penalty_floor(const DeviceState& S, uint32_t ba, uint32_t bb, const StepParams& P) {
  float mA = S.posLin[ba].w, mB = S.posLin[bb].w;
  float big = 3.402823466e38f;
  float minMass = fminf(mA > 0.0f ? mA : big, mB > 0.0f ? mB : big);
  if (minMass == big) return v3s(kPenaltyMin);
  float k = fminf(kPenaltyMax, fmaxf(kPenaltyMin, minMass / fmaxf(P.dt * P.dt, 1e-12f)));
  return v3(k, fminf(k, kPenaltyMaxT), fminf(k, kPenaltyMaxT));
}
__device__ __forceinline__ V3 clamp3(V3 p, V3 lo, V3 hi) { return vmin(vmax(p, lo), hi); }

__device__ bool np_test_axis(const NPBox& A, const NPBox& B, V3 delta, V3 axis, int type, int ia, int ib, V3 relVel,
                             float specCap, const StepParams& P, NPBest& best) {
  float lenSq = dot(axis, axis);
  if (lenSq < kSatEps) return true;
  V3 n = axis * rsq(lenSq);
  if (dot(n, delta) < 0.0f) n = -n;
  float dist = fabsf(dot(delta, n));
  float rA = A.h.x * fabsf(dot(n, A.ax[0])) + A.h.y * fabsf(dot(n, A.ax[1])) + A.h.z * fabsf(dot(n, A.ax[2]));
  float rB = B.h.x * fabsf(dot(n, B.ax[0])) + B.h.y * fabsf(dot(n, B.ax[1])) + B.h.z * fabsf(dot(n, B.ax[2]));
  float sep = dist - (rA + rB);
  float approach = fmaxf(0.0f, -dot(relVel, n));
  float allowed = fminf(approach * P.dt, specCap);
  if (sep > allowed) return false;
  if (!best.valid || sep > best.sep) {
    best.valid = true;
    best.type = type;
    best.ia = ia;
    best.ib = ib;
    best.sep = sep;
    best.n = n;
  }
  return true;
}

__device__ __forceinline__ void np_face_axes(const NPBox& b, int k, V3& u, V3& v, float& eu, float& ev) {
  if (k == 0) { u = b.ax[1]; v = b.ax[2]; eu = b.h.y; ev = b.h.z; }
  else if (k == 1) { u = b.ax[0]; v = b.ax[2]; eu = b.h.x; ev = b.h.z; }
  else { u = b.ax[0]; v = b.ax[1]; eu = b.h.x; ev = b.h.y; }
}

__device__ int np_clip(const V3* in, int inCount, V3 n, float off, V3* out) {
  if (inCount <= 0) return 0;
  int outCount = 0;
  V3 a = in[inCount - 1];
  float da = dot(n, a) - off;
  for (int i = 0; i < inCount; i++) {
    V3 b = in[i];
    float db = dot(n, b) - off;
    bool aIn = da <= kPlaneEps, bIn = db <= kPlaneEps;
    if (aIn != bIn) {
      float t = 0.0f, denom = da - db;
      if (fabsf(denom) > kSatEps) t = clampf(da / denom, 0.0f, 1.0f);
      if (outCount < kMaxPoly) out[outCount++] = a + (b - a) * t;
    }
    if (bIn && outCount < kMaxPoly) out[outCount++] = b;
    a = b;
    da = db;
  }
  return outCount;
}

__device__ void np_support_edge(const NPBox& b, int k, V3 dir, V3& e0, V3& e1) {
  int a1 = (k + 1) % 3, a2 = (k + 2) % 3;
  float s1 = dot(dir, b.ax[a1]) >= 0.0f ? 1.0f : -1.0f;
  float s2 = dot(dir, b.ax[a2]) >= 0.0f ? 1.0f : -1.0f;
  V3 c = b.c + b.ax[a1] * (np_half(b, a1) * s1) + b.ax[a2] * (np_half(b, a2) * s2);
  e0 = c - b.ax[k] * np_half(b, k);
  e1 = c + b.ax[k] * np_half(b, k);
}

// one convex-hull vertex against an oriented box (30_narrowphase.metal:
// 189-247): inside, the nearest exit face; outside, the clamped point within
// the velocity-aware detection margin
struct NPHullCand {
  V3 xHull, xBox, boxToHull;
  float sep;
  uint32_t feature;
};
__device__ __forceinline__ bool np_hull_vertex(V3 xw, uint32_t vid, const NPBox& box, F4 qBox, V3 rel, float hullR,
                                               float boxR, const StepParams& P, NPHullCand& out) {
  const V3 local = q_rotate(q_conj(qBox), xw - box.c);
  const V3 cl = clamp3(local, -box.h, box.h);
  V3 nL, bL;
  float sep;
  const bool inside = local.x >= -box.h.x && local.y >= -box.h.y && local.z >= -box.h.z && local.x <= box.h.x &&
                      local.y <= box.h.y && local.z <= box.h.z;
  if (inside) {
    V3 clr = box.h - vabs(local);
    int axis = 0;
    if (clr.y < comp(clr, axis)) axis = 1;
    if (clr.z < comp(clr, axis)) axis = 2;
    nL = v3s(0);
    setcomp(nL, axis, comp(local, axis) >= 0.0f ? 1.0f : -1.0f);
    bL = local;
    setcomp(bL, axis, comp(nL, axis) * comp(box.h, axis));
    sep = -comp(clr, axis);
  } else {
    V3 d = local - cl;
    float dist = length(d);
    if (dist < 1e-8f) return false;
    nL = d / dist;
    if (dist > np_detect_margin(q_rotate(qBox, nL), rel, P, hullR, boxR)) return false;
    bL = cl;
    sep = dist;
  }
  out.xHull = xw;
  out.xBox = q_rotate(qBox, bL) + box.c;
  out.boxToHull = q_rotate(qBox, nL);
  out.sep = sep;
  out.feature = (4u << 24) | (vid & 0x00FFFFFFu);
  return true;
}

__device__ __forceinline__ V3 np_closest_on_segment(V3 p, V3 a, V3 b) {
  V3 ab = b - a;
  float t = clampf(dot(p - a, ab) / fmaxf(dot(ab, ab), 1e-12f), 0.0f, 1.0f);
  return a + ab * t;
}

__device__ void np_closest_seg_seg(V3 p0, V3 p1, V3 q0, V3 q1, V3& c0, V3& c1) {
  V3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  float s = 0.0f, t = 0.0f;
  if (a <= kSatEps && e <= kSatEps) { c0 = p0; c1 = q0; return; }
  if (a <= kSatEps) {
    t = clampf(f / e, 0.0f, 1.0f);
  } else {
    float c = dot(d1, r);
    if (e <= kSatEps) {
      s = clampf(-c / a, 0.0f, 1.0f);
    } else {
      float b = dot(d1, d2);
      float denom = a * e - b * b;
      if (fabsf(denom) > kSatEps) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
      t = (b * s + f) / e;
      if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
      else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
    }
  }
  c0 = p0 + d1 * s;
  c1 = q0 + d2 * t;
}

__device__ __forceinline__ bool np_add(NPContact* cs, int& count, V3* mids, V3 xA, V3 xB, uint32_t feature) {
  V3 mid = (xA + xB) * 0.5f;
  for (int i = 0; i < count; i++) {
    V3 d = mid - mids[i];
    if (dot(d, d) < kMergeDistSq) return false;
  }
  if (count >= kMaxContacts) return false;
  cs[count].xA = xA;
  cs[count].xB = xB;
  cs[count].feature = feature;
  mids[count] = mid;
  count++;
  return true;
}

__global__ void __launch_bounds__(64) k_np_collide(DeviceState S, StepParams P) {
  uint32_t numPairs = S.counters[CTR_PAIRS];
  for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < numPairs; gid += gridDim.x * blockDim.x) {
    U2 pr = S.pairs[gid];
    uint32_t ia = pr.x, ib = pr.y;
    uint32_t ba = S.colliderOwner[ia], bb = S.colliderOwner[ib];
    F4 bodyPA = S.posLin[ba], bodyPB = S.posLin[bb];
    F4 qBodyA = S.posAng[ba], qBodyB = S.posAng[bb];
    V3 centerA = xyz(bodyPA) + q_rotate(qBodyA, xyz(S.colliderLocalPos[ia]));
    V3 centerB = xyz(bodyPB) + q_rotate(qBodyB, xyz(S.colliderLocalPos[ib]));
    F4 qA = q_mul(qBodyA, S.colliderLocalRot[ia]);
    F4 qB = q_mul(qBodyB, S.colliderLocalRot[ib]);
    V3 vA = xyz(S.velLin[ba]) + cross(xyz(S.velAng[ba]), centerA - xyz(bodyPA));
    V3 vB = xyz(S.velLin[bb]) + cross(xyz(S.velAng[bb]), centerB - xyz(bodyPB));
    V3 relVel = vB - vA;
    F4 shA = S.colliderShape[ia], shB = S.colliderShape[ib];

    NPBox A, B;
    A.c = centerA;
    A.h = xyz(shA) * 0.5f;
    A.ax[0] = q_rotate(qA, v3(1, 0, 0)); A.ax[1] = q_rotate(qA, v3(0, 1, 0)); A.ax[2] = q_rotate(qA, v3(0, 0, 1));
    B.c = centerB;
    B.h = xyz(shB) * 0.5f;
    B.ax[0] = q_rotate(qB, v3(1, 0, 0)); B.ax[1] = q_rotate(qB, v3(0, 1, 0)); B.ax[2] = q_rotate(qB, v3(0, 0, 1));
    V3 delta = B.c - A.c;

    ManifoldRec& outM = S.manifolds[gid];
    outM.colliderPair = U4{ia, ib, 0, 0};
    uint32_t tA = S.colliderShapeType[ia], tB = S.colliderShapeType[ib];
    bool sphA = (tA & kShapeKindMask) == 1, sphB = (tB & kShapeKindMask) == 1;
    F4 fA = S.colliderFriction[ia], fB = S.colliderFriction[ib];
    float staticFriction = combine_friction(fA.x, fB.x, P.frictionCombineMode);
    float dynamicFriction = combine_friction(fA.y, fB.y, P.frictionCombineMode);
    float warm = P.alpha * P.gamma;

    NPContact cs[kMaxContacts];
    int count = 0;
    V3 nrm;  // from B toward A
    const bool capA = (tA & kShapeKindMask) == 3, capB = (tB & kShapeKindMask) == 3;
    const bool hullA = (tA & kShapeKindMask) == 4, hullB = (tB & kShapeKindMask) == 4;
    // convex hulls collide with boxes only (30_narrowphase.metal:404-410);
    // other hull pairs are refused, not approximated by the hull's box
    if ((hullA || hullB) && ((hullA && hullB) || ((hullA ? tB : tA) & kShapeKindMask) != 0)) {
      outM.header = U4{ba, bb, 0, 0};
      continue;
    }

    if (capA || capB || sphA || sphB) {
      float rA = shA.x * 0.5f, rB = shB.x * 0.5f;
      float bestDRound;  // proximity bound of the warm-start match
      if (capA || capB) {
        // capsule pairs (30_narrowphase.metal:510-745): vs sphere and vs
        // capsule one closest-point contact, vs box an alternating
        // clamp <-> segment projection from three seeds
        const bool cIsA = capA;
        const F4 shC = cIsA ? shA : shB, shO = cIsA ? shB : shA;
        const F4 qC = cIsA ? qA : qB, qO = cIsA ? qB : qA;
        const V3 pC = cIsA ? A.c : B.c, pO = cIsA ? B.c : A.c;
        const float rc = shC.y;
        const V3 axC = q_rotate(qC, v3(0, 0, 1));
        const V3 c0 = pC - axC * (shC.x * 0.5f), c1 = pC + axC * (shC.x * 0.5f);
        const uint32_t otherType = (cIsA ? tB : tA) & kShapeKindMask;
        const V3 relCO = cIsA ? (vB - vA) : (vA - vB);
        V3 nCO = v3(0, 0, 1);  // capsule -> other
        NPContact hits[3];     // xA on the capsule, xB on the other shape
        int nh = 0;
        if (otherType == 1) {
          const float rs = shO.x * 0.5f;
          V3 q = np_closest_on_segment(pO, c0, c1);
          V3 d = pO - q;
          float dist = length(d);
          V3 nT = dist > 1e-6f ? d / dist : v3(0, 0, 1);
          if (dist <= rc + rs + np_detect_margin(nT, relCO, P, shC.w, shO.w)) {
            nCO = nT;
            hits[0].xA = q + nCO * rc;
            hits[0].xB = pO - nCO * rs;
            hits[0].feature = 0u;
            nh = 1;
          }
        } else if (otherType == 3) {
          const float ro = shO.y;
          const V3 axO = q_rotate(qO, v3(0, 0, 1));
          const V3 o0 = pO - axO * (shO.x * 0.5f), o1 = pO + axO * (shO.x * 0.5f);
          V3 qa, qb;
          np_closest_seg_seg(c0, c1, o0, o1, qa, qb);
          V3 d = qb - qa;
          float dist = length(d);
          V3 nT = dist > 1e-6f ? d / dist : v3(0, 0, 1);
          if (dist <= rc + ro + np_detect_margin(nT, relCO, P, shC.w, shO.w)) {
            nCO = nT;
            hits[0].xA = qa + nCO * rc;
            hits[0].xB = qb - nCO * ro;
            hits[0].feature = 0u;
            nh = 1;
          }
        } else {
          const F4 qOc = q_conj(qO);
          const V3 halfO = xyz(shO) * 0.5f;
          V3 nAccum = v3s(0);
          for (int seed = 0; seed < 3 && nh < 3; seed++) {
            V3 q = c0 + (c1 - c0) * ((float)seed * 0.5f);
            V3 bw = q;
            for (int k = 0; k < 5; k++) {
              V3 lb = clamp3(q_rotate(qOc, q - pO), -halfO, halfO);
              bw = q_rotate(qO, lb) + pO;
              q = np_closest_on_segment(bw, c0, c1);
            }
            V3 d = bw - q;
            float dist = length(d);
            V3 n;
            if (dist < 1e-6f) {
              // segment inside the box: push out along the least-penetrated face
              V3 lb = q_rotate(qOc, q - pO);
              V3 pen = halfO - vabs(lb);
              int axis = 0;
              if (pen.y < comp(pen, axis)) axis = 1;
              if (pen.z < comp(pen, axis)) axis = 2;
              V3 nl = v3s(0);
              setcomp(nl, axis, comp(lb, axis) >= 0.0f ? 1.0f : -1.0f);
              n = -q_rotate(qO, nl);
              bw = q - n * comp(pen, axis);
            } else {
              n = d / dist;
              if (dist - rc > np_detect_margin(n, relCO, P, shC.w, shO.w)) continue;
            }
            bool dup = false;
            for (int h = 0; h < nh; h++)
              if (length(hits[h].xA - (q + n * rc)) < 0.4f * rc + 0.05f) { dup = true; break; }
            if (dup) continue;
            hits[nh].xA = q + n * rc;
            hits[nh].xB = bw;
            hits[nh].feature = (uint32_t)seed;
            nAccum += n;
            nh++;
          }
          if (nh > 0) nCO = normalize(nAccum);
        }
        if (nh == 0) { outM.header = U4{ba, bb, 0, 0}; continue; }
        nrm = cIsA ? -nCO : nCO;
        for (int h = 0; h < nh; h++) {
          cs[h].xA = cIsA ? hits[h].xA : hits[h].xB;
          cs[h].xB = cIsA ? hits[h].xB : hits[h].xA;
          cs[h].feature = hits[h].feature;
        }
        count = nh;
        bestDRound = 0.6f * (rc + 0.2f);
      } else if (sphA && sphB) {
        V3 d = A.c - B.c;
        float dist = length(d);
        V3 nT = dist > 1e-9f ? d / dist : v3(0, 0, 1);
        float detectM = np_detect_margin(nT, vA - vB, P, rA, rB);
        if (dist > rA + rB + detectM) { outM.header = U4{ba, bb, 0, 0}; continue; }
        nrm = nT;
        cs[0].xA = A.c - nrm * rA;
        cs[0].xB = B.c + nrm * rB;
        cs[0].feature = 0;
        count = 1;
      } else {
        bool sIsA = sphA;
        float r = sIsA ? rA : rB;
        const NPBox& box = sIsA ? B : A;
        F4 qBox = sIsA ? qB : qA;
        V3 sc = sIsA ? A.c : B.c;
        V3 local = q_rotate(q_conj(qBox), sc - box.c);
        V3 half = box.h;
        V3 clamped = clamp3(local, -half, half);
        V3 nLocal, qLocal;
        if (clamped.x == local.x && clamped.y == local.y && clamped.z == local.z) {
          V3 pen = half - vabs(local);
          int axis = 0;
          if (pen.y < comp(pen, axis)) axis = 1;
          if (pen.z < comp(pen, axis)) axis = 2;
          V3 n = v3s(0);
          setcomp(n, axis, comp(local, axis) >= 0.0f ? 1.0f : -1.0f);
          nLocal = n;
          qLocal = local;
          setcomp(qLocal, axis, comp(n, axis) * comp(half, axis));
        } else {
          V3 d = local - clamped;
          float dist = length(d);
          V3 nLT = d / fmaxf(dist, 1e-9f);
          V3 nWT = q_rotate(qBox, nLT);
          V3 vS = sIsA ? vA : vB, vBx = sIsA ? vB : vA;
          float detectM = np_detect_margin(nWT, vS - vBx, P, r, sIsA ? shB.w : shA.w);
          if (dist > r + detectM) { outM.header = U4{ba, bb, 0, 0}; continue; }
          nLocal = nLT;
          qLocal = clamped;
        }
        V3 nW = q_rotate(qBox, nLocal);
        V3 qW = q_rotate(qBox, qLocal) + box.c;
        V3 xS = sc - nW * r;
        nrm = sIsA ? nW : -nW;
        cs[0].xA = sIsA ? xS : qW;
        cs[0].xB = sIsA ? qW : xS;
        cs[0].feature = 0;
        count = 1;
      }
      if (!(capA || capB)) bestDRound = 0.5f * fmaxf(rA, rB);
      // round-shape tail: anchors stored as world offsets when flagged
      V3 t1, t2;
      orthonormal(nrm, t1, t2);
      int prevIdx = pair_map_find(S, P.mapCapacity, ia, ib);
      bool roundA = (tA & kShapeWorldRound) != 0, roundB = (tB & kShapeWorldRound) != 0;
      uint32_t flags = 1u | (roundA ? 2u : 0u) | (roundB ? 4u : 0u);
      outM.header = U4{ba, bb, (uint32_t)count, flags};
      outM.basisN = f4(nrm, dynamicFriction);
      outM.basisT1 = f4(t1, staticFriction);
      F4 qAc = q_conj(qBodyA), qBc = q_conj(qBodyB);
      for (int i = 0; i < count; i++) {
        V3 rA_ = roundA ? (cs[i].xA - xyz(bodyPA)) : q_rotate(qAc, cs[i].xA - xyz(bodyPA));
        V3 rB_ = roundB ? (cs[i].xB - xyz(bodyPB)) : q_rotate(qBc, cs[i].xB - xyz(bodyPB));
        V3 lambda = v3s(0), penalty = v3s(0);
        float stick = 0.0f;
        if (prevIdx >= 0) {
          const ManifoldRec& pm = S.prevManifolds[prevIdx];
          uint32_t pn = pm.header.z;
          float bestD = bestDRound;
          int bestJ = -1;
          for (uint32_t j = 0; j < pn; j++) {
            float dj = length(xyz(pm.c[j].rA) - rA_);
            if (dj < bestD) { bestD = dj; bestJ = (int)j; }
          }
          if (bestJ >= 0) {
            lambda = xyz(pm.c[bestJ].lambda);
            penalty = xyz(pm.c[bestJ].penalty);
            V3 t1o = xyz(pm.basisT1);
            V3 t2o = cross(xyz(pm.basisN), t1o);
            V3 lt = t1o * lambda.y + t2o * lambda.z;
            lambda.y = dot(t1, lt);
            lambda.z = dot(t2, lt);
            // body-local anchors inside the Coulomb cone are kept; a
            // world-offset (rolling) anchor of a rigid round shape is
            // rebuilt -- capsules keep anchors only when neither is one
            bool partA = shA.w < 0.0f, partB = shB.w < 0.0f;
            bool keep = (capA || capB) ? (!roundA && !roundB) : ((partA || !roundA) && (partB || !roundB));
            if (keep) {
              stick = pm.c[bestJ].rB.w;
              if (stick != 0.0f) {
                rA_ = xyz(pm.c[bestJ].rA);
                rB_ = xyz(pm.c[bestJ].rB);
              }
            }
          }
        }
        V3 xAw = roundA ? xyz(bodyPA) + rA_ : xform(xyz(bodyPA), qBodyA, rA_);
        V3 xBw = roundB ? xyz(bodyPB) + rB_ : xform(xyz(bodyPB), qBodyB, rB_);
        V3 d = xAw - xBw;
        V3 C0 = v3(dot(nrm, d) + P.collisionMargin, dot(t1, d), dot(t2, d));
        lambda *= warm;
        V3 fl = np_penalty_floor(S, ba, bb, P);
        penalty = v3(clampf(penalty.x * P.gamma, fl.x, kPenaltyMax), clampf(penalty.y * P.gamma, fl.y, kPenaltyMaxT),
                     clampf(penalty.z * P.gamma, fl.z, kPenaltyMaxT));
        outM.c[i].rA = f4(rA_, as_float(cs[i].feature));
        outM.c[i].rB = f4(rB_, stick);
        outM.c[i].C0 = f4(C0, 0.0f);
        outM.c[i].lambda = f4(lambda, 0.0f);
        outM.c[i].penalty = f4(penalty, 0.0f);
      }
      continue;
    }

    if (hullA || hullB) {
      // convex hull vs box (30_narrowphase.metal:411-497): the deepest hull
      // vertex, then up to kMaxContacts - 1 more within 0.75 margin of it
      // along the same face normal, each maximizing its least tangential
      // distance to those chosen (ties to the lower vertex id)
      const bool hIsA = hullA;
      const uint32_t ih = hIsA ? ia : ib;
      const F4 qH = hIsA ? qA : qB, qBox = hIsA ? qB : qA;
      const V3 pH = hIsA ? A.c : B.c;
      const NPBox& box = hIsA ? B : A;
      const V3 rel = hIsA ? (vA - vB) : (vB - vA);
      const float hullR = fabsf(hIsA ? shA.w : shB.w), boxR = fabsf(hIsA ? shB.w : shA.w);
      const U2 range = S.colliderHullRange[ih];
      NPHullCand best, chosen[kMaxContacts];
      bool haveBest = false;
      for (uint32_t v = 0; v < range.y; v++) {
        NPHullCand c;
        if (!np_hull_vertex(pH + q_rotate(qH, xyz(S.hullVerts[range.x + v])), v, box, qBox, rel, hullR, boxR, P, c))
          continue;
        if (!haveBest || c.sep < best.sep - 1e-8f || (fabsf(c.sep - best.sep) <= 1e-8f && c.feature < best.feature)) {
          best = c;
          haveBest = true;
        }
      }
      if (!haveBest) { outM.header = U4{ba, bb, 0, 0}; continue; }
      chosen[0] = best;
      int nc = 1;
      const float band = 0.75f * P.collisionMargin;
      while (nc < kMaxContacts) {
        NPHullCand next;
        bool haveNext = false;
        float bestSpread = -1.0f;
        for (uint32_t v = 0; v < range.y; v++) {
          NPHullCand c;
          if (!np_hull_vertex(pH + q_rotate(qH, xyz(S.hullVerts[range.x + v])), v, box, qBox, rel, hullR, boxR, P, c))
            continue;
          if (c.sep > best.sep + band || dot(c.boxToHull, best.boxToHull) < 0.98f) continue;
          bool dup = false;
          float minSpread = 3.402823466e38f;
          for (int j = 0; j < nc; j++) {
            if (c.feature == chosen[j].feature) { dup = true; break; }
            V3 d = c.xHull - chosen[j].xHull;
            d = d - best.boxToHull * dot(best.boxToHull, d);
            minSpread = fminf(minSpread, dot(d, d));
          }
          if (dup) continue;
          if (!haveNext || minSpread > bestSpread + 1e-10f ||
              (fabsf(minSpread - bestSpread) <= 1e-10f && c.feature < next.feature)) {
            next = c;
            bestSpread = minSpread;
            haveNext = true;
          }
        }
        if (!haveNext || bestSpread < kMergeDistSq) break;
        chosen[nc++] = next;
      }
      for (int i = 0; i < nc; i++) {
        cs[i].xA = hIsA ? chosen[i].xHull : chosen[i].xBox;
        cs[i].xB = hIsA ? chosen[i].xBox : chosen[i].xHull;
        cs[i].feature = chosen[i].feature;
      }
      count = nc;
      nrm = hIsA ? best.boxToHull : -best.boxToHull;
    } else {
    // ---- box-box SAT ----
    NPBest bestFace, bestEdge;
    bestFace.valid = false; bestFace.sep = -3.402823466e38f;
    bestEdge.valid = false; bestEdge.sep = -3.402823466e38f;
    float cap = np_spec_cap(shA.w, shB.w, P.collisionMargin);
    bool separated = false;
    for (int i = 0; i < 3 && !separated; i++)
      if (!np_test_axis(A, B, delta, A.ax[i], 0, i, -1, relVel, cap, P, bestFace)) separated = true;
    for (int i = 0; i < 3 && !separated; i++)
      if (!np_test_axis(A, B, delta, B.ax[i], 1, -1, i, relVel, cap, P, bestFace)) separated = true;
    for (int i = 0; i < 3 && !separated; i++)
      for (int j = 0; j < 3 && !separated; j++)
        if (!np_test_axis(A, B, delta, cross(A.ax[i], B.ax[j]), 2, i, j, relVel, cap, P, bestEdge)) separated = true;
    if (separated || !bestFace.valid) { outM.header = U4{ba, bb, 0, 0}; continue; }
    NPBest best = bestFace;
    if (bestEdge.valid && 0.95f * bestEdge.sep > bestFace.sep + 0.01f) best = bestEdge;

    V3 mids[kMaxContacts];
    if (best.type == 2) {
      V3 a0, a1, b0, b1, xA, xB;
      np_support_edge(A, best.ia, best.n, a0, a1);
      np_support_edge(B, best.ib, -best.n, b0, b1);
      np_closest_seg_seg(a0, a1, b0, b1, xA, xB);
      uint32_t feature = (2u << 24) | ((uint32_t)(best.ia & 0xFF) << 8) | (uint32_t)(best.ib & 0xFF);
      np_add(cs, count, mids, xA, xB, feature);
      if (count == 0) np_add(cs, count, mids, np_support(A, best.n), np_support(B, -best.n), feature);
    } else {
      bool refIsA = best.type == 0;
      int refAxis = refIsA ? best.ia : best.ib;
      const NPBox& ref = refIsA ? A : B;
      const NPBox& inc = refIsA ? B : A;
      V3 refOut = refIsA ? best.n : -best.n;
      float signR = dot(refOut, ref.ax[refAxis]) >= 0.0f ? 1.0f : -1.0f;
      V3 refN = ref.ax[refAxis] * signR;
      V3 refC = ref.c + refN * np_half(ref, refAxis);
      V3 refU, refV;
      float refEU, refEV;
      np_face_axes(ref, refAxis, refU, refV, refEU, refEV);
      int incAxis = 0;
      float bestD = -3.402823466e38f;
      for (int i = 0; i < 3; i++) {
        float d = fabsf(dot(inc.ax[i], refN));
        if (d > bestD) { bestD = d; incAxis = i; }
      }
      float signI = dot(inc.ax[incAxis], refN) > 0.0f ? -1.0f : 1.0f;
      V3 incN = inc.ax[incAxis] * signI;
      V3 incC = inc.c + incN * np_half(inc, incAxis);
      V3 incU, incV;
      float incEU, incEV;
      np_face_axes(inc, incAxis, incU, incV, incEU, incEV);
      V3 p0[kMaxPoly], p1[kMaxPoly];
      p0[0] = incC + incU * incEU + incV * incEV;
      p0[1] = incC - incU * incEU + incV * incEV;
      p0[2] = incC - incU * incEU - incV * incEV;
      p0[3] = incC + incU * incEU - incV * incEV;
      int n = 4;
      n = np_clip(p0, n, refU, dot(refU, refC) + refEU, p1);
      n = np_clip(p1, n, -refU, dot(-refU, refC) + refEU, p0);
      n = np_clip(p0, n, refV, dot(refV, refC) + refEV, p1);
      n = np_clip(p1, n, -refV, dot(-refV, refC) + refEV, p0);
      uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | ((uint32_t)(refAxis & 0xFF) << 16) |
                        ((uint32_t)(incAxis & 0xFF) << 8);
      float slop = kPlaneEps + fmaxf(0.0f, best.sep);
      for (int i = 0; i < n && count < kMaxContacts; i++) {
        V3 pInc = p0[i];
        float d = dot(pInc - refC, refN);
        if (d > slop) continue;
        V3 pRef = pInc - refN * d;
        np_add(cs, count, mids, refIsA ? pRef : pInc, refIsA ? pInc : pRef, prefix | (uint32_t)(i & 0xFF));
      }
      if (count == 0) np_add(cs, count, mids, np_support(A, best.n), np_support(B, -best.n), prefix);
    }
    nrm = -best.n;
    }  // box-box SAT
    V3 t1, t2;
    orthonormal(nrm, t1, t2);
    int prevIdx = pair_map_find(S, P.mapCapacity, ia, ib);
    outM.header = U4{ba, bb, (uint32_t)count, 1u};
    outM.basisN = f4(nrm, dynamicFriction);
    outM.basisT1 = f4(t1, staticFriction);
    F4 qAc = q_conj(qBodyA), qBc = q_conj(qBodyB);
    V3 fl = np_penalty_floor(S, ba, bb, P);
    for (int i = 0; i < count; i++) {
      V3 rA = q_rotate(qAc, cs[i].xA - xyz(bodyPA));
      V3 rB = q_rotate(qBc, cs[i].xB - xyz(bodyPB));
      V3 lambda = v3s(0), penalty = v3s(0);
      float stick = 0.0f;
      if (prevIdx >= 0) {
        const ManifoldRec& pm = S.prevManifolds[prevIdx];
        uint32_t pn = pm.header.z;
        for (uint32_t j = 0; j < pn; j++) {
          if (as_uint(pm.c[j].rA.w) == cs[i].feature) {
            lambda = xyz(pm.c[j].lambda);
            penalty = xyz(pm.c[j].penalty);
            stick = pm.c[j].rB.w;
            if (stick != 0.0f) {
              rA = xyz(pm.c[j].rA);
              rB = xyz(pm.c[j].rB);
            }
            break;
          }
        }
      }
      V3 xAw = xform(xyz(bodyPA), qBodyA, rA);
      V3 xBw = xform(xyz(bodyPB), qBodyB, rB);
      V3 d = xAw - xBw;
      V3 C0 = v3(dot(nrm, d) + P.collisionMargin, dot(t1, d), dot(t2, d));
      lambda *= warm;
      penalty = v3(clampf(penalty.x * P.gamma, fl.x, kPenaltyMax), clampf(penalty.y * P.gamma, fl.y, kPenaltyMaxT),
                   clampf(penalty.z * P.gamma, fl.z, kPenaltyMaxT));
      outM.c[i].rA = f4(rA, as_float(cs[i].feature));
      outM.c[i].rB = f4(rB, stick);
      outM.c[i].C0 = f4(C0, 0.0f);
      outM.c[i].lambda = f4(lambda, 0.0f);
      outM.c[i].penalty = f4(penalty, 0.0f);
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(uint32_t n, uint32_t tpb) { return (n + tpb - 1) / tpb; }

void launch_broadphase(hipStream_t st, const DeviceState& S, const StepParams& P) {
  const bool fine = P.fineCellSize > 0.0f && P.numHashed > 0;
  if (P.numHashed > 0) {
    hipLaunchKernelGGL(k_bp_count, dim3(blocks_for(P.numHashed, 256)), dim3(256), 0, st, S, P);
  }
  if (fine) hipLaunchKernelGGL(k_bp_fine_count, dim3(blocks_for(P.numHashed, 256)), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.cellCount, S.cellStart, P.gridHashSize, S.scanScratch);
  if (fine) launch_exclusive_scan(st, S.fineCount, S.fineStart, P.gridHashSize, S.scanScratch);
  if (P.numHashed > 0) {
    hipLaunchKernelGGL(k_bp_scatter, dim3(blocks_for(P.numHashed, 256)), dim3(256), 0, st, S, P);
  }
  if (fine) hipLaunchKernelGGL(k_bp_fine_scatter, dim3(blocks_for(P.numHashed, 256)), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_bp_sort_cells, dim3(blocks_for(P.gridHashSize, 256)), dim3(256), 0, st, S, P);
  {
    uint32_t large = P.numHashed / 33u + 1u;
    hipLaunchKernelGGL(k_bp_sort_large, dim3(large < 1024u ? large : 1024u), dim3(256), 0, st, S, P);
  }
  if (P.numHashed > 0) {
    hipLaunchKernelGGL(k_bp_gather, dim3(blocks_for(P.numHashed, 256)), dim3(256), 0, st, S, P);
  }
  uint32_t producers = P.numHashed + P.numGlobals;
  if (producers > 0) {
    hipLaunchKernelGGL(k_bp_pairs<false>, dim3(blocks_for(producers, 4)), dim3(256), 0, st, S, P);
    launch_exclusive_scan(st, S.pairCount, S.pairStart, producers, S.scanScratch);
    hipLaunchKernelGGL(k_bp_pairs<true>, dim3(blocks_for(producers, 4)), dim3(256), 0, st, S, P);
  }
  hipLaunchKernelGGL(k_bp_finalize, dim3(1), dim3(64), 0, st, S, P);
}

void launch_narrowphase(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  uint32_t blocks = blocks_for(P.maxPairs, 64);
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_np_collide, dim3(blocks), dim3(64), 0, st, S, P);
}

void launch_map_rebuild(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  launch_fill_u32(st, S.mapKeyA, P.mapCapacity, 0u);
  uint32_t blocks = blocks_for(P.maxPairs, 256);
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_pm_insert, dim3(blocks), dim3(256), 0, st, S, P);
}

void launch_map_find(hipStream_t st, const DeviceState& S, const StepParams& P, const U2* q, int32_t* out,
                     uint32_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_map_find, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, P, q, out, n);
}

}  // namespace avbd
