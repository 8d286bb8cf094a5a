// collide.hip -- spatial hash, deterministic rigid pair stream and CAS
// persistence map for gfx950 (the narrow phase is narrow.hip).
//
// Integer outputs of this file (hash cells, cell lists, pair sets, map
// contents) must be bit-reproducible, so floating-point contraction is off
// here: every product/sum rounds exactly as written.
#include "kernels.hpp"
#include "grid.hpp"
#include "pairmap.hpp"

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
__device__ __forceinline__ void bp_producer(const DeviceState& S, const StepParams& P, uint32_t producer,
                                            uint64_t (*fineKeys)[2][kFineCap]) {
  const uint32_t lane = threadIdx.x & 63u;
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

// The wave path: one wave per producer of bpWave[0, bpWaveCount) -- the
// producers not on the lane list (globals, hashed producers of small dense
// packings), then those the lane path deferred this frame.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_bp_pairs(DeviceState S, StepParams P) {
  __shared__ uint64_t fineKeys[4][2][kFineCap];
  const uint32_t n = S.bpWaveCount[0];
  for (uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6); w < n; w += gridDim.x * 4)
    bp_producer<EMIT>(S, P, S.bpWave[w], fineKeys);
}

// Hashed producers on the lane path, one lane each, with the wave path's
// filters and order: the 27 neighbour cells in (dz, dy, dx) order (a hash met
// earlier in that order is not walked again), each bucket in collider order,
// then the globals.  A soft-group particle with no hashed non-particle
// collider in the scene skips the cells.  Under the candidate search grid the
// lane gathers its accepted partners from the 27 search cells, keys each by
// (position of its hash bucket in the visit order, collider id) and keeps up
// to kLaneCap of them sorted in registers -- the walk's set and order, as on
// the wave path.  A producer with more than kLaneCandidates candidates in its
// cells, or more than kLaneCap accepted search-grid partners, goes to the
// wave path: the count pass appends it to bpWave (one atomic per wave), the
// emit pass reaches the same decision and leaves it there.  The lane list
// (built at create) holds every hashed producer without the search grid or
// when there are many of them (one lane each hides the walk's latency better
// than one wave each), else the soft-group particles that scan only globals.
constexpr uint32_t kLaneCap = 8;
constexpr uint32_t kLaneCandidates = 64;

__device__ __forceinline__ void bp_defer(const DeviceState& S, bool defer, uint32_t producer) {
  const uint64_t m = __ballot(defer);
  if (m == 0ull) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)__activemask()) - 1);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(S.bpWaveCount, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader, 64);
  if (defer) S.bpWave[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = producer;
}

template <bool EMIT>
__global__ void __launch_bounds__(256) k_bp_lane(DeviceState S, StepParams P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.numBpLane) return;
  const uint32_t producer = S.bpLane[i];
  const uint32_t a = S.hashedIdx[producer];
  const U4 ma = S.hashMeta[producer];
  const F4 ca = S.hashCenter[producer];
  const uint32_t ownerA = ma.x, grpA = ma.y, shA = ma.z & 1u, ga = ma.w;
  const bool aDyn = (ma.z & 2u) != 0;
  const V3 pa = xyz(ca);
  const float ra = ca.w;
  const bool cells = !(ga != 0 && P.numHashedRigid == 0);
  const bool fine = P.fineCellSize > 0.0f;
  // the 27 cells walked (hash cells, or search cells under the search grid)
  // and their ranges, duplicates and skipped cells empty
  uint32_t h[27], cs[27], cc[27];
  uint32_t total = 0;
  int cx = 0, cy = 0, cz = 0;
  if (cells) {
    cell_coord(pa, P.cellSize, cx, cy, cz);
#pragma unroll
    for (int q = 0; q < 27; q++)
      h[q] = cell_hash(cx + q % 3 - 1, cy + (q / 3) % 3 - 1, cz + q / 9 - 1, grpA, P.gridHashSize);
    int wx = cx, wy = cy, wz = cz;
    if (fine) cell_coord(pa, P.fineCellSize, wx, wy, wz);
#pragma unroll
    for (int q = 0; q < 27; q++) {
      const uint32_t hq =
          fine ? cell_hash(wx + q % 3 - 1, wy + (q / 3) % 3 - 1, wz + q / 9 - 1, grpA, P.gridHashSize) : h[q];
      bool dup = false;
#pragma unroll
      for (int j = 0; j < q; j++)
        dup |= (fine ? cell_hash(wx + j % 3 - 1, wy + (j / 3) % 3 - 1, wz + j / 9 - 1, grpA, P.gridHashSize) : h[j]) ==
               hq;
      const bool skip = dup || (!fine && ga != 0 && S.cellRigid[hq] == 0);
      cs[q] = skip ? 0u : (fine ? S.fineStart[hq] : S.cellStart[hq]);
      cc[q] = skip ? 0u : (fine ? S.fineCount[hq] : S.cellCount[hq]);
      total += cc[q];
    }
  }
  bool defer = total > kLaneCandidates;
  uint64_t keys[kLaneCap];
  uint32_t nAcc = 0;
  if (!defer && cells && fine) {
#pragma unroll
    for (uint32_t k = 0; k < kLaneCap; k++) keys[k] = ~0ull;
#pragma unroll
    for (int q = 0; q < 27; q++) {
      for (uint32_t k = cs[q]; k < cs[q] + cc[q]; k++) {
        const uint32_t g = S.fineBodies[k];
        const uint32_t b = S.hashedIdx[g];
        if (b <= a) continue;
        const U4 mb = S.hashMeta[g];
        const F4 cb = S.hashCenter[g];
        if ((ga != 0 && mb.w != 0) ||
            !pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, mb.x, mb.y, mb.z & 1u, (mb.z & 2u) != 0, xyz(cb),
                         cb.w))
          continue;
        int bx, by, bz;
        cell_coord(xyz(cb), P.cellSize, bx, by, bz);
        const uint32_t hb = cell_hash(bx, by, bz, S.colliderGroup[b], P.gridHashSize);
        if (ga != 0 && S.cellRigid[hb] == 0) continue;
        uint32_t pos = 27;
#pragma unroll
        for (int j = 26; j >= 0; j--) pos = h[j] == hb ? (uint32_t)j : pos;
        if (pos == 27) continue;
        // sorted insertion (ascending keys; static register indices)
        uint64_t x = ((uint64_t)pos << 32) | b;
#pragma unroll
        for (uint32_t m = 0; m < kLaneCap; m++) {
          const uint64_t lo = keys[m] < x ? keys[m] : x, hi = keys[m] < x ? x : keys[m];
          keys[m] = lo;
          x = hi;
        }
        nAcc++;
      }
    }
    defer = nAcc > kLaneCap;
  }
  if (!EMIT) bp_defer(S, defer, producer);
  if (defer) return;
  const uint32_t base = EMIT ? S.pairStart[producer] : 0u;
  uint32_t count = 0;
  auto emit = [&](uint32_t x, uint32_t y) {
    if (EMIT && base + count < P.maxPairs) S.pairs[base + count] = U2{x, y};
    count++;
  };
  if (cells && fine) {
#pragma unroll
    for (uint32_t k = 0; k < kLaneCap; k++)
      if (k < nAcc) emit(a, (uint32_t)(keys[k] & 0xFFFFFFFFull));
  } else if (cells) {
#pragma unroll
    for (int q = 0; q < 27; q++) {
      for (uint32_t k = cs[q]; k < cs[q] + cc[q]; k++) {
        const uint32_t b = S.cellBodies[k];
        if (b <= a) continue;
        const U4 mb = S.cellMeta[k];
        const F4 cb = S.cellPos[k];
        if (!(ga != 0 && mb.w != 0) &&
            pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, mb.x, mb.y, mb.z & 1u, (mb.z & 2u) != 0, xyz(cb),
                        cb.w))
          emit(a, b);
      }
    }
  }
  for (uint32_t g = 0; g < P.numGlobals; g++) {
    const uint32_t b = S.globalIdx[g];
    const uint32_t ownerB = S.colliderOwner[b];
    if (!(ga != 0 && S.clothGroup[ownerB] == ga) &&
        pair_accept(S, P, ownerA, grpA, shA, aDyn, pa, ra, ownerB, S.colliderGroup[b], S.colliderShared[b],
                    S.posLin[ownerB].w > 0.0f, collider_center(S, b), fabsf(S.colliderShape[b].w)))
      emit(a < b ? a : b, a < b ? b : a);
  }
  if (!EMIT) S.pairCount[producer] = count;
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

__global__ void k_map_find(DeviceState S, StepParams P, const U2* q, int32_t* out, uint32_t n) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  out[gid] = pair_map_find(S, P.mapCapacity, q[gid].x, q[gid].y);
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
    // the wave path's grid: its static list plus a bounded share for
    // deferred lane producers (the kernel loops over its list)
    const uint32_t deferred = blocks_for(P.numBpLane, 4);
    const uint32_t waveBlocks = blocks_for(P.numBpWave, 4) + (deferred < 1024u ? deferred : 1024u);
    launch_fill_u32(st, S.bpWaveCount, 1u, P.numBpWave);
    if (P.numBpLane > 0)
      hipLaunchKernelGGL(k_bp_lane<false>, dim3(blocks_for(P.numBpLane, 256)), dim3(256), 0, st, S, P);
    if (waveBlocks > 0) hipLaunchKernelGGL(k_bp_pairs<false>, dim3(waveBlocks), dim3(256), 0, st, S, P);
    launch_exclusive_scan(st, S.pairCount, S.pairStart, producers, S.scanScratch);
    if (P.numBpLane > 0)
      hipLaunchKernelGGL(k_bp_lane<true>, dim3(blocks_for(P.numBpLane, 256)), dim3(256), 0, st, S, P);
    if (waveBlocks > 0) hipLaunchKernelGGL(k_bp_pairs<true>, dim3(waveBlocks), dim3(256), 0, st, S, P);
  }
  hipLaunchKernelGGL(k_bp_finalize, dim3(1), dim3(64), 0, st, S, P);
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
