// solve.hip -- AVBD solve kernels for gfx950: inertial predict, contact
// adjacency, static / contact-aware coloring, lane-split primal block
// descent, augmented-Lagrangian dual, BDF1 finalize, residual, diagnostics,
// and the persistent cooperative sweep.
#include "kernels.hpp"
#include "stamps.hpp"

namespace avbd {

__device__ __forceinline__ bool body_dynamic(const DeviceState& S, uint32_t b) {
  return b != kWorld && S.posLin[b].w > 0.0f;
}

// Grid barrier for cooperative launches (every workgroup co-resident).
// Every wait is bounded: past the spin bound a block records CTR_GBAR_FAIL
// and proceeds, and once any block has given up later barriers do not wait,
// so the grid always drains.
// One release (cache write-back) before arriving, relaxed polling of the
// arrival counter, and one acquire (cache invalidate) once it is reached:
// an acquire load inside the spin would invalidate the L2 on every poll.
__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t* fail, uint32_t target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target &&
           __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 20)) {
        atomicAdd(fail, 1u);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// per-frame contact adjacency: count / scan / scatter / sort of active
// manifolds into a second CSR beside the static topology CSR
// (40_solver.metal:18-243 restricted to the per-frame forces)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cadj_count(DeviceState S, StepParams P) {
  uint32_t n = S.counters[CTR_PAIRS], ns = S.counters[CTR_SOFT];
  uint32_t active = 0;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n + ns; m += gridDim.x * blockDim.x) {
    if (m >= n) {
      // element contact: one entry per dynamic participant
      U4 ids = S.soft[m - n].ids;
      uint32_t v[4] = {ids.x, ids.y, ids.z, ids.w};
      for (int k = 0; k < 4; k++)
        if (body_dynamic(S, v[k])) atomicAdd(&S.dAdjCount[v[k]], 1u);
      continue;
    }
    U4 h = S.manifolds[m].header;
    if (h.z == 0) continue;
    active++;
    if (body_dynamic(S, h.x)) atomicAdd(&S.dAdjCount[h.x], 1u);
    if (body_dynamic(S, h.y)) atomicAdd(&S.dAdjCount[h.y], 1u);
  }
  if (active) atomicAdd(&S.counters[CTR_ACTIVE_MANIFOLDS], active);
}
__global__ void __launch_bounds__(256) k_cadj_cursor(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numBodies) return;
  S.dAdjCursor[gid] = S.dAdjStart[gid];
}
__global__ void __launch_bounds__(256) k_cadj_scatter(DeviceState S, StepParams P) {
  uint32_t n = S.counters[CTR_PAIRS], ns = S.counters[CTR_SOFT];
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n + ns; m += gridDim.x * blockDim.x) {
    if (m >= n) {
      // same-color participants (static palette) are read from the snapshot
      uint32_t si = m - n;
      U4 ids = S.soft[si].ids;
      uint32_t v[4] = {ids.x, ids.y, ids.z, ids.w};
      bool dv[4];
      for (int k = 0; k < 4; k++) dv[k] = body_dynamic(S, v[k]);
      for (int k = 0; k < 4; k++) {
        if (!dv[k]) continue;
        uint32_t same = 0;
        if (P.staticPalette) {
          uint32_t ck = S.colors[v[k]];
          for (int j = 0; j < 4; j++)
            if (j != k && dv[j] && v[j] != v[k] && S.colors[v[j]] == ck) same |= 1u << j;
        }
        uint32_t entry = (FK_SOFT << kAdjShift) | (same << kSoftSameShift) | si;
        S.dAdjList[atomicAdd(&S.dAdjCursor[v[k]], 1u)] = entry;
      }
      continue;
    }
    U4 h = S.manifolds[m].header;
    if (h.z == 0) continue;
    uint32_t entry = (FK_MANIFOLD << kAdjShift) | m;
    bool da = body_dynamic(S, h.x), db = body_dynamic(S, h.y);
    if (da) S.dAdjList[atomicAdd(&S.dAdjCursor[h.x], 1u)] = entry;
    if (db) S.dAdjList[atomicAdd(&S.dAdjCursor[h.y], 1u)] = entry;
    if (P.staticPalette && da && db && S.colors[h.x] == S.colors[h.y])
      S.manifolds[m].header.w = h.w | kManifoldSameColor;
  }
}
// Sorting each (short) segment makes the accumulation order a function of
// the manifold indices only, which are themselves deterministic.
__global__ void __launch_bounds__(256) k_cadj_sort(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numBodies) return;
  uint32_t s = S.dAdjStart[gid], n = S.dAdjCount[gid];
  for (uint32_t i = 1; i < n; i++) {
    uint32_t v = S.dAdjList[s + i];
    uint32_t j = i;
    while (j > 0 && S.dAdjList[s + j - 1] > v) {
      S.dAdjList[s + j] = S.dAdjList[s + j - 1];
      j--;
    }
    S.dAdjList[s + j] = v;
  }
}

// dAdjCount must be zero on entry (cleared with the frame's counters)
void launch_contact_adjacency(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  uint32_t blocks = (P.maxPairs + P.maxSoft + 255) / 256;
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_cadj_count, dim3(blocks), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.dAdjCount, S.dAdjStart, P.numBodies, S.scanScratch);
  uint32_t bb = (P.numBodies + 255) / 256;
  hipLaunchKernelGGL(k_cadj_cursor, dim3(bb), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_cadj_scatter, dim3(blocks), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_cadj_sort, dim3(bb), dim3(256), 0, st, S, P);
}

// ---------------------------------------------------------------------------
// neighbour enumeration for coloring: the other dynamic participants of an
// adjacency entry (40_solver.metal:260-318)
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ void for_each_neighbor(const DeviceState& S, uint32_t entry, uint32_t self, bool staticOnly,
                                                  F&& f) {
  uint32_t kind = entry >> kAdjShift, idx = entry & kAdjIndexMask;
  if (kind == FK_TET) {
    U4 ids = S.tets[idx].ids;
    uint32_t v[4] = {ids.x, ids.y, ids.z, ids.w};
    for (int k = 0; k < 4; k++)
      if (v[k] != self) f(v[k]);
  } else if (kind == FK_JOINT) {
    U4 h = S.joints[idx].header;
    f(h.x == self ? h.y : h.x);
  } else if (kind == FK_SPRING || kind == FK_PSPRING) {
    U4 h = S.springs[idx].header;
    f(h.x == self ? h.y : h.x);
  } else if (kind == FK_MANIFOLD && !staticOnly) {
    U4 h = S.manifolds[idx].header;
    f(h.x == self ? h.y : h.x);
  }
}

// Static palette: c_v = lowest color unused by lower-index dynamic clique
// neighbours (GPUSolver.swift:1956-2050).  Jacobi iteration of this rule
// converges to the serial greedy result: once every lower neighbour holds
// its final color, v's color is final, so the iteration settles within the
// longest increasing-index path of the graph.  64 = palette exhausted.
__global__ void __launch_bounds__(256) k_static_color(DeviceState S, StepParams P, const uint32_t* in, uint32_t* out) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.numBodies) return;
  if (!body_dynamic(S, v)) { out[v] = 0; return; }
  uint64_t used = 0;
  uint32_t s = S.sAdjStart[v], n = S.sAdjCount[v];
  for (uint32_t k = 0; k < n; k++) {
    for_each_neighbor(S, S.sAdjList[s + k], v, true, [&](uint32_t nb) {
      if (nb == kWorld || nb >= v || !body_dynamic(S, nb)) return;
      uint32_t c = in[nb];
      if (c < 64) used |= (1ull << c);
    });
  }
  uint32_t c = used == ~0ull ? 64u : (uint32_t)__builtin_ctzll(~used);
  out[v] = c;
  if (c != in[v]) S.counters[CTR_COLOR_CHANGED] = 1u;
}

void launch_static_color_pass(hipStream_t st, const DeviceState& S, const StepParams& P, const uint32_t* in,
                              uint32_t* out) {
  hipLaunchKernelGGL(k_static_color, dim3((P.numBodies + 255) / 256), dim3(256), 0, st, S, P, in, out);
}

// ---------------------------------------------------------------------------
// contact-aware dynamic coloring for rigid-only scenes (40_solver.metal:
// 320-540): 20 Jacobi passes (smaller-index conflict rule + compaction),
// conditional serial greedy repair, validation, stable color buckets.
// ---------------------------------------------------------------------------
struct ColorMasks {
  uint64_t lower, all;
  bool conflict;
};
__device__ __forceinline__ ColorMasks neighbor_masks(const DeviceState& S, uint32_t v, const uint32_t* colors,
                                                     uint32_t myColor) {
  ColorMasks m{0ull, 0ull, false};
  auto visit = [&](uint32_t nb) {
    if (nb == kWorld || !body_dynamic(S, nb)) return;
    uint32_t c = colors[nb];
    uint64_t bit = c < 64 ? (1ull << c) : 0ull;
    m.all |= bit;
    if (nb < v) {
      m.lower |= bit;
      if (c == myColor) m.conflict = true;
    }
  };
  uint32_t s = S.sAdjStart[v], n = S.sAdjCount[v];
  for (uint32_t k = 0; k < n; k++) for_each_neighbor(S, S.sAdjList[s + k], v, false, visit);
  s = S.dAdjStart[v];
  n = S.dAdjCount[v];
  for (uint32_t k = 0; k < n; k++) for_each_neighbor(S, S.dAdjList[s + k], v, false, visit);
  return m;
}

__global__ void __launch_bounds__(256) k_color_iterate(DeviceState S, StepParams P, const uint32_t* in, uint32_t* out,
                                                      uint32_t pass) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.numBodies) return;
  if (pass > 0 && S.colorFlags[pass - 1] == 0u) return;  // converged: dst already holds the fixed point
  if (S.posLin[v].w <= 0.0f) { out[v] = 0; return; }
  uint32_t my = in[v];
  ColorMasks m = neighbor_masks(S, v, in, my);
  if (m.conflict) {
    uint64_t freeL = ~m.lower;
    uint32_t nc = freeL ? (uint32_t)__builtin_ctzll(freeL) : 63u;
    out[v] = nc > 63u ? 63u : nc;
    S.colorFlags[pass] = 1u;
  } else {
    uint64_t freeA = ~m.all;
    uint32_t best = freeA ? (uint32_t)__builtin_ctzll(freeA) : my;
    out[v] = best < my ? best : my;
    if (best < my) S.colorFlags[pass] = 1u;
  }
}

// Repair of an unfinished parallel coloring (40_solver.metal:387-436).  The
// reference walks all bodies serially, giving each the lowest color unused
// by its lower-index neighbours.  That serial result is the unique fixed
// point of the same rule applied in any order: a body's value is final once
// all lower-index neighbours are final.  Each workgroup owns a contiguous
// range of bodies and re-applies the rule to it until the range stops
// changing (workgroup barriers only: chains inside the range resolve
// there), then the grid meets at a barrier; a round in which no workgroup's
// first local pass changed anything has reached the fixed point.  Chains
// that cross ranges cost a grid round each; in index-ordered scenes they are
// rare.  Cooperative launch; flags F[21 + r % 3] count changing workgroups
// of round r.
__device__ __forceinline__ bool repair_pass(const DeviceState& S, uint32_t* colors, uint32_t lo, uint32_t hi) {
  bool changed = false;
  for (uint32_t v = lo + threadIdx.x; v < hi; v += blockDim.x) {
    uint32_t cur = __hip_atomic_load(&colors[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t nc = 0;
    if (S.posLin[v].w > 0.0f) {
      ColorMasks m = neighbor_masks(S, v, colors, cur);
      uint64_t freeL = ~m.lower;
      nc = freeL ? (uint32_t)__builtin_ctzll(freeL) : 63u;
    }
    if (nc != cur) {
      __hip_atomic_store(&colors[v], nc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      changed = true;
    }
  }
  return __syncthreads_or(changed) != 0;
}

__global__ void __launch_bounds__(256) k_color_repair_coop(DeviceState S, StepParams P, uint32_t* colors,
                                                           uint32_t finalPass) {
  if (S.colorFlags[finalPass] == 0u) return;  // grid-uniform: the 20 Jacobi passes converged
  uint32_t* bar = &S.counters[CTR_REPAIR_BAR];
  uint32_t* fail = &S.counters[CTR_GBAR_FAIL];
  uint32_t epoch = 0;
  const uint32_t chunk = (P.numBodies + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = blockIdx.x * chunk < P.numBodies ? blockIdx.x * chunk : P.numBodies;
  const uint32_t hi = lo + chunk < P.numBodies ? lo + chunk : P.numBodies;
  for (uint32_t round = 0; round <= P.numBodies; round++) {
    if (blockIdx.x == 0 && threadIdx.x == 0) S.colorFlags[21 + (round + 1) % 3] = 0u;
    // local passes: the first one's change is what this round reports
    bool first = repair_pass(S, colors, lo, hi);
    if (first) {
      for (uint32_t k = 1; k <= hi - lo && repair_pass(S, colors, lo, hi); k++) {
      }
      if (threadIdx.x == 0) atomicAdd(&S.colorFlags[21 + round % 3], 1u);
    }
    epoch += gridDim.x;
    grid_barrier(bar, fail, epoch);
    uint32_t any = __hip_atomic_load(&S.colorFlags[21 + round % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (any == 0u || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
      if (blockIdx.x == 0 && threadIdx.x == 0) S.counters[CTR_REPAIR_PASSES] = round + 1;
      return;
    }
  }
}

// Serial form, used only when cooperative launch is unavailable.
__global__ void k_color_repair(DeviceState S, StepParams P, uint32_t* colors, uint32_t finalPass) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (S.colorFlags[finalPass] == 0u) return;
  for (uint32_t v = 0; v < P.numBodies; v++) {
    if (S.posLin[v].w <= 0.0f) { colors[v] = 0; continue; }
    ColorMasks m = neighbor_masks(S, v, colors, colors[v]);
    uint64_t freeL = ~m.lower;
    colors[v] = freeL ? (uint32_t)__builtin_ctzll(freeL) : 63u;
  }
}

static int repair_coop_blocks(uint32_t numBodies) {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, cus = 0, per = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_color_repair_coop, 256, 0) != hipSuccess) per = 0;
    cached = per * cus;
  }
  int want = (int)((numBodies + 255) / 256);
  return want < cached ? want : cached;
}

__global__ void __launch_bounds__(256) k_color_validate(DeviceState S, StepParams P, const uint32_t* colors) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.numBodies || S.posLin[v].w <= 0.0f) return;
  ColorMasks m = neighbor_masks(S, v, colors, colors[v]);
  if (m.conflict) atomicAdd(&S.counters[CTR_COLOR_CONFLICTS], 1u);
}

// Stable bucket build: per-block color histograms (block-major within each
// color after the scan) and an in-order scatter; ranks inside a wave come
// from ballots, so colorList is sorted by body index within each color.
__global__ void __launch_bounds__(256) k_color_hist(DeviceState S, StepParams P, const uint32_t* colors,
                                                   uint32_t* hist, uint32_t numBlocks) {
  __shared__ uint32_t h[kMaxColors];
  if (threadIdx.x < kMaxColors) h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < P.numBodies && S.posLin[v].w > 0.0f) atomicAdd(&h[colors[v] & 63u], 1u);
  __syncthreads();
  if (threadIdx.x < kMaxColors) {
    hist[threadIdx.x * numBlocks + blockIdx.x] = h[threadIdx.x];
    if (h[threadIdx.x]) atomicAdd(&S.counters[CTR_COLOR_BASE + threadIdx.x], h[threadIdx.x]);
  }
}
__global__ void __launch_bounds__(256) k_color_scatter(DeviceState S, StepParams P, const uint32_t* colors,
                                                      const uint32_t* histStart, uint32_t numBlocks) {
  __shared__ uint32_t waveCount[4][kMaxColors];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < 4 * kMaxColors; i += blockDim.x) (&waveCount[0][0])[i] = 0;
  __syncthreads();
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  bool dyn = v < P.numBodies && S.posLin[v].w > 0.0f;
  uint32_t c = dyn ? (colors[v] & 63u) : 0u;
  uint32_t rank = 0;
  uint64_t pending = __ballot(dyn);
  const uint64_t ltMask = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  while (pending) {
    uint32_t leader = (uint32_t)__builtin_ctzll(pending);
    uint32_t lc = __shfl(c, leader, 64);
    uint64_t same = __ballot(dyn && c == lc);
    if (dyn && c == lc) rank = (uint32_t)__builtin_popcountll(same & ltMask);
    if (lane == 0) waveCount[wave][lc] = (uint32_t)__builtin_popcountll(same);
    pending &= ~same;
  }
  __syncthreads();
  if (dyn) {
    uint32_t off = histStart[c * numBlocks + blockIdx.x];
    for (uint32_t w = 0; w < wave; w++) off += waveCount[w][c];
    S.colorList[off + rank] = v;
  }
}
__global__ void k_color_starts(DeviceState S, StepParams P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t run = 0, used = 0;
  for (uint32_t c = 0; c < kMaxColors; c++) {
    S.colorStart[c] = run;
    uint32_t n = S.counters[CTR_COLOR_BASE + c];
    run += n;
    if (n > 0) used = c + 1;
  }
  S.colorStart[kMaxColors] = run;
  S.colorStart[kMaxColors + 1] = used;
}

void launch_dynamic_coloring(hipStream_t st, const DeviceState& S, const StepParams& P) {
  const uint32_t passes = 20;
  uint32_t bb = (P.numBodies + 255) / 256;
  launch_fill_u32(st, S.colorFlags, passes + 4, 0u);
  uint32_t* src = S.colors;
  uint32_t* dst = S.colorsB;
  for (uint32_t p = 0; p < passes; p++) {
    hipLaunchKernelGGL(k_color_iterate, dim3(bb), dim3(256), 0, st, S, P, src, dst, p);
    uint32_t* t = src; src = dst; dst = t;
  }
  // after an even number of passes the final colors are back in S.colors
  int rb = repair_coop_blocks(P.numBodies);
  bool coop = false;
  if (rb > 0) {
    DeviceState s2 = S;
    StepParams p2 = P;
    uint32_t* c2 = src;
    uint32_t fp = passes - 1;
    void* args[] = {&s2, &p2, &c2, &fp};
    coop = hipLaunchCooperativeKernel((const void*)k_color_repair_coop, dim3(rb), dim3(256), args, 0, st) ==
           hipSuccess;
    if (!coop) (void)hipGetLastError();
  }
  if (!coop) hipLaunchKernelGGL(k_color_repair, dim3(1), dim3(64), 0, st, S, P, src, passes - 1);
  hipLaunchKernelGGL(k_color_validate, dim3(bb), dim3(256), 0, st, S, P, src);
  // buckets: hist lives in the scan scratch tail of dAdjCursor-sized scratch
  uint32_t* hist = S.bodySlot;  // needs 64 * bb words; sized at create
  uint32_t* histStart = S.bodySlot + kMaxColors * bb;
  hipLaunchKernelGGL(k_color_hist, dim3(bb), dim3(256), 0, st, S, P, src, hist, bb);
  launch_exclusive_scan(st, hist, histStart, kMaxColors * bb, S.scanScratch);
  hipLaunchKernelGGL(k_color_scatter, dim3(bb), dim3(256), 0, st, S, P, src, histStart, bb);
  hipLaunchKernelGGL(k_color_starts, dim3(1), dim3(64), 0, st, S, P);
}

// ---------------------------------------------------------------------------
// lane-split primal: L lanes stamp one body's adjacency slice each, DPP /
// ds_swizzle butterflies reduce the block system, lane 0 solves
// (40_solver.metal:1544-1685 re-mapped onto wave64).
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float swz16_f(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}
// Sum over aligned groups of L lanes; every lane of a group gets the total.
template <int L>
__device__ __forceinline__ float group_sum(float v) {
  if (L >= 2) v += dpp_f<0xB1>(v);    // quad_perm [1,0,3,2]: xor 1
  if (L >= 4) v += dpp_f<0x4E>(v);    // quad_perm [2,3,0,1]: xor 2
  if (L >= 8) v += dpp_f<0x141>(v);   // row_half_mirror: other quad of 8
  if (L >= 16) v += dpp_f<0x140>(v);  // row_mirror: other half of 16
  if (L >= 32) v += swz16_f(v);       // ds_swizzle xor 16
  return v;
}
template <int L>
__device__ __forceinline__ V3 group_sum3(V3 v) {
  return v3(group_sum<L>(v.x), group_sum<L>(v.y), group_sum<L>(v.z));
}

// Isotropic limiter (40_solver.metal:1361-1382): a particle's total
// displacement from the pose its bound holds around stays within
// max(0.45 sqrt(bound), 0.2 r); `cut` reports a truncation (counted for the
// refresh check and the frame's diagnostics by ogc_count).
__device__ __forceinline__ V3 ogc_truncate(const DeviceState& S, uint32_t body, V3 cur, V3 dx, float rad, bool& cut) {
  const float d2 = as_float(S.ogcBound[body]);
  if (!(d2 < 1e37f)) return dx;
  const float bv = fmaxf(0.45f * sqrtf(d2), 0.2f * rad);
  const V3 anchor = xyz(S.ogcPrev[body]);
  V3 tot = (cur + dx) - anchor;
  const float t2 = dot(tot, tot);
  if (t2 > bv * bv) {
    tot = tot * (bv / sqrtf(t2));
    dx = (anchor + tot) - cur;
    cut = true;
  }
  return dx;
}
// The wave's truncations in one pair of atomics (the lanes still running
// here are the solving lanes of the wave's bodies): per-body atomics on
// the two shared counters serialize when most particles are held back.
__device__ __forceinline__ void ogc_count(const DeviceState& S, bool cut) {
  const uint64_t m = __ballot(cut);
  if (m != 0ull && __lane_id() == (uint32_t)(__ffsll((unsigned long long)__activemask()) - 1)) {
    const uint32_t n = (uint32_t)__popcll(m);
    atomicAdd(&S.counters[CTR_OGC], n);
    atomicAdd(&S.counters[CTR_OGC_TRUNC], n);
  }
}

// Gathers and stamps one body's entries (lane-strided), reduces, and lane 0
// solves and writes.  Must be reached by whole waves (EXEC full) because of
// the ballot and DPP reductions; inactive groups pass active = false.
template <int L, int MODE>
__device__ __forceinline__ void primal_group(const DeviceState& S, const StepParams& P, bool active, U4 work,
                                             uint32_t lane, uint32_t color) {
  constexpr bool SOFT = MODE == kModeTet;
  const uint32_t body = work.x;
  bool rigid = MODE == kModeGeneral && active && S.shape[body].w >= 0.0f;
  const bool anyRigid = MODE == kModeGeneral && __ballot(rigid) != 0ull;
  Acc acc;
  acc_zero(acc);
  if (active) {
    PoseView pv{S.posLin, S.posAng, S.posLinSnap, S.posAngSnap, true};
    const V3 xs = xyz(S.posLin[body]), ls = xyz(S.posLo[body]);
    const uint32_t ss = work.y, sc = work.z;
    const uint32_t ds = S.dAdjStart[body], dc = S.dAdjCount[body];
    if (SOFT) {
      // particle colors: static entries are tets, per-frame entries manifolds.
      // Each lane first issues all of its slot-record loads of a batch, then
      // all position gathers, then the arithmetic.
      constexpr int B = L >= 32 ? 1 : (L == 16 ? 2 : 4);
      for (uint32_t base = lane; base < sc; base += L * B) {
        TetSlot r[B];
#pragma unroll
        for (int i = 0; i < B; i++) {
          uint32_t k = base + i * L;
          if (k < sc) r[i] = S.tetSlots[ss + k];
        }
        V3 x[B][3], lo[B][3];
#pragma unroll
        for (int i = 0; i < B; i++) {
          if (base + i * L < sc) {
            x[i][0] = xyz(S.posLin[r[i].other.x]);
            x[i][1] = xyz(S.posLin[r[i].other.y]);
            x[i][2] = xyz(S.posLin[r[i].other.z]);
            lo[i][0] = xyz(S.posLo[r[i].other.x]);
            lo[i][1] = xyz(S.posLo[r[i].other.y]);
            lo[i][2] = xyz(S.posLo[r[i].other.z]);
          }
        }
#pragma unroll
        for (int i = 0; i < B; i++)
          if (base + i * L < sc)
            stamp_tet_slot_e(r[i], split_edge(x[i][0], lo[i][0], xs, ls), split_edge(x[i][1], lo[i][1], xs, ls),
                             split_edge(x[i][2], lo[i][2], xs, ls), acc);
      }
      // contact entries continue the lane round-robin after the tets
      for (uint32_t k = (lane + L - (sc % L)) % L; k < dc; k += L) {
        uint32_t entry = S.dAdjList[ds + k];
        if ((entry >> kAdjShift) == FK_SOFT)
          stamp_soft(S.soft[entry & kSoftIndexMask], body, (entry >> kSoftSameShift) & 15u, pv, S.initLin, S.initAng,
                     P.alpha, acc);
        else
          stamp_manifold(S.manifolds[entry & kAdjIndexMask], body, false, pv, S.initLin, S.initAng, P.alpha, acc, ls);
      }
    } else if (MODE == kModeParticle) {
      // particles: cloth slots inline, any other static entry generically
      for (uint32_t k = lane; k < sc; k += L) {
        uint32_t entry = S.sAdjList[ss + k];
        uint32_t kind = entry >> kAdjShift;
        if (kind >= FK_MEMBRANE) stamp_cloth_slot(kind, S.clothSlots[ss + k], pv, xs, S.springs, P.alpha, acc);
        else stamp_entry(S, pv, entry, ss + k, body, xs, ls, false, P.alpha, P.dt, acc);
      }
      for (uint32_t k = (lane + L - (sc % L)) % L; k < dc; k += L) {
        uint32_t entry = S.dAdjList[ds + k];
        if ((entry >> kAdjShift) == FK_SOFT)
          stamp_soft(S.soft[entry & kSoftIndexMask], body, (entry >> kSoftSameShift) & 15u, pv, S.initLin, S.initAng,
                     P.alpha, acc);
        else
          stamp_manifold(S.manifolds[entry & kAdjIndexMask], body, false, pv, S.initLin, S.initAng, P.alpha, acc, ls);
      }
    } else {
      // static entries round-robin over the lanes; each contact manifold
      // split over the lanes contact by contact (a body's few manifolds
      // carry up to 8 contacts each), other per-frame entries round-robin
      for (uint32_t k = lane; k < sc; k += L) stamp_entry(S, pv, S.sAdjList[ss + k], ss + k, body, xs, ls, rigid,
                                                           P.alpha, P.dt, acc);
      for (uint32_t j = 0; j < dc; j++) {
        const uint32_t entry = S.dAdjList[ds + j];
        if ((entry >> kAdjShift) == FK_MANIFOLD)
          stamp_manifold(S.manifolds[entry & kAdjIndexMask], body, rigid, pv, S.initLin, S.initAng, P.alpha, acc, ls,
                         lane, (uint32_t)L);
        else if ((sc + j) % L == lane)
          stamp_entry(S, pv, entry, ss + sc + j, body, xs, ls, rigid, P.alpha, P.dt, acc);
      }
    }
  }
  acc.l00 = group_sum<L>(acc.l00);
  acc.l10 = group_sum<L>(acc.l10);
  acc.l11 = group_sum<L>(acc.l11);
  acc.l20 = group_sum<L>(acc.l20);
  acc.l21 = group_sum<L>(acc.l21);
  acc.l22 = group_sum<L>(acc.l22);
  acc.gl = group_sum3<L>(acc.gl);
  if (anyRigid) {
    acc.a00 = group_sum<L>(acc.a00);
    acc.a10 = group_sum<L>(acc.a10);
    acc.a11 = group_sum<L>(acc.a11);
    acc.a20 = group_sum<L>(acc.a20);
    acc.a21 = group_sum<L>(acc.a21);
    acc.a22 = group_sum<L>(acc.a22);
    acc.cr.r0 = group_sum3<L>(acc.cr.r0);
    acc.cr.r1 = group_sum3<L>(acc.cr.r1);
    acc.cr.r2 = group_sum3<L>(acc.cr.r2);
    acc.ga = group_sum3<L>(acc.ga);
  }
  if (!active || lane != 0) return;
  F4 pl = S.posLin[body];
  const V3 lo = xyz(S.posLo[body]);
  float mass = pl.w, dt2 = P.dt * P.dt;
  float md = mass / dt2;
  acc.l00 += md; acc.l11 += md; acc.l22 += md;
  acc.gl += ((xyz(pl) - xyz(S.inertLin[body])) + lo) * md;
  V3 dxL, dxA = v3s(0);
  float rad = fabsf(S.shape[body].w);
  if (rigid) {
    F4 q = S.posAng[body];
    M3 inertia = m3_scale(world_inertia(q, xyz(S.props[body])), 1.0f / dt2);
    acc_ang_m3(acc, inertia);
    acc.ga += m3_mul(inertia, q_sub(q, S.inertAng[body]));
    solve6x6(lower_to_m3(acc.l00, acc.l10, acc.l11, acc.l20, acc.l21, acc.l22),
             lower_to_m3(acc.a00, acc.a10, acc.a11, acc.a20, acc.a21, acc.a22), acc.cr, -acc.gl, -acc.ga, dxL, dxA);
    if (!finite3(dxL) || !finite3(dxA)) return;
    float a2 = dot(dxA, dxA);
    if (a2 > 0.25f) dxA *= 0.5f * rsq(a2);
  } else {
    dxL = solve3x3(acc.l00, acc.l10, acc.l11, acc.l20, acc.l21, acc.l22, -acc.gl);
    if (!finite3(dxL)) return;
  }
  float maxLin = 0.35f * rad;
  float l2 = dot(dxL, dxL);
  if (l2 > maxLin * maxLin) dxL *= maxLin * rsq(l2);
  if (P.isoOn) {
    bool cut = false;
    if (!rigid) dxL = ogc_truncate(S, body, xyz(pl), dxL, rad, cut);
    ogc_count(S, cut);
  }
  // (hi, lo) + dx, renormalised exactly: hi = f32(hi + t), lo = the error term
  V3 hiN, loN;
  two_sum3(xyz(pl), lo + dxL, hiN, loN);
  S.posLin[body] = f4(hiN, mass);
  S.posLo[body] = f4(loN, 0.0f);
  if (rigid) S.posAng[body] = q_addw(S.posAng[body], dxA);
}

// (body, static CSR start, static CSR count) of a body of a device-built palette
__device__ __forceinline__ U4 body_work(const DeviceState& S, bool active, uint32_t body) {
  return active ? U4{body, S.sAdjStart[body], S.sAdjCount[body], 0u} : U4{0u, 0u, 0u, 0u};
}

// XCD-aware block order: the hardware deals consecutive workgroups to the
// 8 XCDs round-robin; remapping gives each XCD one contiguous slice of the
// color list so neighbour gathers of adjacent bodies share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t X = 8;
  uint32_t per = (n + X - 1) / X;
  uint32_t r = (b % X) * per + b / X;
  return r;
}

template <int L, int MODE>
__global__ void __launch_bounds__(256) k_primal(DeviceState S, StepParams P, uint32_t color, uint32_t count,
                                                uint32_t base) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;  // the frame's residual check has ended its iterations
  const uint32_t groupsPerBlock = 256 / L;
  uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  uint32_t slot = blk * groupsPerBlock + threadIdx.x / L;
  uint32_t lane = threadIdx.x & (L - 1);
  U4 work;
  bool active;
  if (count == 0xFFFFFFFFu) {
    // device-sized (dynamic palette): color list, then the body's CSR range
    uint32_t s = S.colorStart[color];
    count = S.colorStart[color + 1] - s;
    active = slot < count;
    work = body_work(S, active, active ? S.colorList[s + slot] : 0u);
  } else {
    // static palette: one load of the precomputed (body, CSR start, count)
    active = slot < count;
    work = active ? S.colorWork[base + slot] : U4{0u, 0u, 0u, 0u};
  }
  primal_group<L, MODE>(S, P, active, work, lane, color);
}

// Colors [first, 64) of a device-sized palette in one workgroup, in order,
// with a workgroup barrier between colors.  Used for the (normally empty)
// tail of a dynamic palette so each iteration launches a fixed, small
// number of kernels; same per-body arithmetic as k_primal, so the result
// does not depend on which kernel processed a color.
template <int L>
__global__ void __launch_bounds__(256) k_primal_tail(DeviceState S, StepParams P, uint32_t first) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;
  const uint32_t groupsPerBlock = 256 / L;
  const uint32_t lane = threadIdx.x & (L - 1), grp = threadIdx.x / L;
  const uint32_t used = S.colorStart[kMaxColors + 1];
  for (uint32_t c = first; c < used; c++) {
    uint32_t s = S.colorStart[c], count = S.colorStart[c + 1] - s;
    for (uint32_t base = 0; base < count; base += groupsPerBlock) {
      uint32_t slot = base + grp;
      bool active = slot < count;
      primal_group<L, kModeGeneral>(S, P, active, body_work(S, active, active ? S.colorList[s + slot] : 0u), lane, c);
    }
    __threadfence_block();
    __syncthreads();
  }
}

void launch_primal_tail(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t first, int lanes) {
  switch (lanes) {
    case 4: hipLaunchKernelGGL(k_primal_tail<4>, dim3(1), dim3(256), 0, st, S, P, first); break;
    case 8: hipLaunchKernelGGL(k_primal_tail<8>, dim3(1), dim3(256), 0, st, S, P, first); break;
    case 16: hipLaunchKernelGGL(k_primal_tail<16>, dim3(1), dim3(256), 0, st, S, P, first); break;
    default: hipLaunchKernelGGL(k_primal_tail<32>, dim3(1), dim3(256), 0, st, S, P, first); break;
  }
}

template <int MODE>
static void launch_primal_t(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color,
                            uint32_t count, uint32_t base, int lanes, uint32_t blocks) {
  switch (lanes) {
    case 4: hipLaunchKernelGGL((k_primal<4, MODE>), dim3(blocks), dim3(256), 0, st, S, P, color, count, base); break;
    case 8: hipLaunchKernelGGL((k_primal<8, MODE>), dim3(blocks), dim3(256), 0, st, S, P, color, count, base); break;
    case 16:
      hipLaunchKernelGGL((k_primal<16, MODE>), dim3(blocks), dim3(256), 0, st, S, P, color, count, base);
      break;
    default:
      hipLaunchKernelGGL((k_primal<32, MODE>), dim3(blocks), dim3(256), 0, st, S, P, color, count, base);
      break;
  }
}

void launch_primal_color(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color, uint32_t count,
                         uint32_t base, int lanes, int mode) {
  if (count == 0) return;
  uint32_t groupsPerBlock = 256 / lanes;
  uint32_t n = count == 0xFFFFFFFFu ? P.numBodies : count;
  uint32_t blocks = (n + groupsPerBlock - 1) / groupsPerBlock;
  // the XCD remap needs the padded slot range to cover every body
  blocks = (blocks + 7) / 8 * 8;
  if (mode == kModeTet) launch_primal_t<kModeTet>(st, S, P, color, count, base, lanes, blocks);
  else if (mode == kModeParticle) launch_primal_t<kModeParticle>(st, S, P, color, count, base, lanes, blocks);
  else launch_primal_t<kModeGeneral>(st, S, P, color, count, base, lanes, blocks);
}

// ---------------------------------------------------------------------------
// augmented-Lagrangian dual + penalty ramp (40_solver.metal:1689-1910)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dual_joint(const DeviceState& S, const StepParams& P, uint32_t gid) {
  JointRec& j = S.joints[gid];
  if (j.header.z != 0) return;
  uint32_t a = j.header.x, b = j.header.y;
  float torqueArm = j.C0Lin.w;
  V3 penLin = xyz(j.penaltyLin);
  if (dot(penLin, penLin) > 0.0f) {
    V3 pA = a == kWorld ? xyz(j.rA) : xform(xyz(S.posLin[a]), S.posAng[a], xyz(j.rA));
    V3 pB = xform(xyz(S.posLin[b]), S.posAng[b], xyz(j.rB));
    V3 C = pA - pB;
    if (j.header.w & 1u) {
      C -= xyz(j.C0Lin) * P.alpha;
      V3 l = penLin * C + xyz(j.lambdaLin);
      j.lambdaLin = f4(v3(clampf(l.x, -P.lambdaMax, P.lambdaMax), clampf(l.y, -P.lambdaMax, P.lambdaMax),
                          clampf(l.z, -P.lambdaMax, P.lambdaMax)),
                       j.lambdaLin.w);
    }
    float cap = fminf(j.rA.w, kPenaltyMax);
    V3 np = penLin + vabs(C) * P.betaLin;
    j.penaltyLin = f4(v3(fminf(np.x, cap), fminf(np.y, cap), fminf(np.z, cap)), j.penaltyLin.w);
  }
  V3 penAng = xyz(j.penaltyAng);
  if (dot(penAng, penAng) > 0.0f) {
    F4 qA = a == kWorld ? F4{0, 0, 0, 1} : S.posAng[a];
    V3 C;
    if (j.hingeAxis.w != 0.0f) {
      V3 aB = q_rotate(S.posAng[b], xyz(j.hingeAxis));
      V3 aA = q_rotate(q_mul(qA, j.restRel), xyz(j.hingeAxis));
      C = cross(aA, aB) * torqueArm;
    } else {
      C = q_sub(q_mul(qA, j.restRel), S.posAng[b]) * torqueArm;
    }
    if (j.header.w & 2u) {
      C -= xyz(j.C0Ang) * P.alpha;
      V3 l = penAng * C + xyz(j.lambdaAng);
      j.lambdaAng = f4(v3(clampf(l.x, -P.lambdaMax, P.lambdaMax), clampf(l.y, -P.lambdaMax, P.lambdaMax),
                          clampf(l.z, -P.lambdaMax, P.lambdaMax)),
                       j.lambdaAng.w);
    }
    float cap = fminf(j.rB.w, kPenaltyMax);
    V3 np = penAng + vabs(C) * P.betaAng;
    j.penaltyAng = f4(v3(fminf(np.x, cap), fminf(np.y, cap), fminf(np.z, cap)), j.penaltyAng.w);
  }
  float fracture = j.C0Ang.w;
  V3 la = xyz(j.lambdaAng);
  float lin2 = (j.header.w & 8u) ? dot(xyz(j.lambdaLin), xyz(j.lambdaLin)) : 0.0f;
  if ((j.header.w & 4u) && dot(la, la) + lin2 > fracture * fracture) {
    j.penaltyLin = F4{0, 0, 0, 0};
    j.penaltyAng = F4{0, 0, 0, 0};
    j.lambdaLin = F4{0, 0, 0, 0};
    j.lambdaAng = F4{0, 0, 0, 0};
    j.header.z = 1;
  }
}

__device__ __forceinline__ void dual_spring(const DeviceState& S, const StepParams& P, uint32_t si) {
  SpringRec& sp = S.springs[si];
  if (sp.header.z == 0) return;
  uint32_t a = sp.header.x, b = sp.header.y;
  V3 pA = xform(xyz(S.posLin[a]), S.posAng[a], xyz(sp.rA));
  V3 pB = xform(xyz(S.posLin[b]), S.posAng[b], xyz(sp.rB));
  float C = length(pA - pB) - sp.rB.w - sp.dual.z * P.alpha;
  float mA = S.posLin[a].w, mB = S.posLin[b].w;
  float big = 3.402823466e38f;
  float mMin = fminf(mA > 0.0f ? mA : big, mB > 0.0f ? mB : big);
  float cap = fminf(P.lambdaMax, fmaxf(2.0f, 5.0e3f * mMin));
  sp.dual.x = clampf(0.98f * sp.dual.x + sp.dual.y * C, 0.0f, cap);
  if (C > 0.0f) sp.dual.y = fminf(sp.dual.y + C * P.betaLin, fminf(sp.rA.w, kPenaltyMax));
}

// One item per contact slot (kMaxContacts per manifold): a contact's update
// reads only its own record and the manifold's shared header, basis and
// body poses, so the slots of a manifold run on adjacent lanes and their
// records load as one contiguous span instead of a serial loop per thread.
__device__ __forceinline__ void dual_contact(const DeviceState& S, const StepParams& P, uint32_t gid, uint32_t i) {
  ManifoldRec& m = S.manifolds[gid];
  uint32_t n = m.header.z;
  if (i >= n) return;
  uint32_t a = m.header.x, b = m.header.y;
  bool rA_round = (m.header.w & 2u) != 0, rB_round = (m.header.w & 4u) != 0;
  V3 dqAL = xyz(S.posLin[a]) - xyz(S.initLin[a]);
  V3 dqAA = q_sub(S.posAng[a], S.initAng[a]);
  V3 dqBL = xyz(S.posLin[b]) - xyz(S.initLin[b]);
  V3 dqBA = q_sub(S.posAng[b], S.initAng[b]);
  float mA = S.posLin[a].w, mB = S.posLin[b].w;
  float big = 3.402823466e38f;
  float minMass = fminf(mA > 0.0f ? mA : big, mB > 0.0f ? mB : big);
  float lamCap = fminf(P.lambdaMax, fmaxf(10.0f, 1.0e5f * minMass));
  V3 nrm = xyz(m.basisN), t1 = xyz(m.basisT1), t2 = cross(nrm, t1);
  {
    V3 rAW = rA_round ? xyz(m.c[i].rA) : q_rotate(S.posAng[a], xyz(m.c[i].rA));
    V3 rBW = rB_round ? xyz(m.c[i].rB) : q_rotate(S.posAng[b], xyz(m.c[i].rB));
    V3 C;
    float fs, bnd;
    V3 F = contact_force(m, (int)i, dqAL, dqAA, dqBL, dqBA, rAW, rBW, P.alpha, nrm, t1, t2, C, fs, bnd);
    F.x = fmaxf(F.x, -lamCap);
    m.c[i].lambda = f4(F, 0.0f);
    V3 pen = xyz(m.c[i].penalty);
    if (F.x < 0.0f) pen.x = fminf(pen.x + P.betaLin * fabsf(C.x), kPenaltyMax);
    if (fs <= bnd) {
      pen.y = fminf(pen.y + P.betaLin * fabsf(C.y), kPenaltyMaxT);
      pen.z = fminf(pen.z + P.betaLin * fabsf(C.z), kPenaltyMaxT);
      m.c[i].rB.w = sqrtf(C.y * C.y + C.z * C.z) < kStickThresh ? 1.0f : 0.0f;
    }
    m.c[i].penalty = f4(pen, 0.0f);
  }
}

// element contact dual: the quadratic stage only (the log-barrier stage is
// dual-free), mass-scaled normal cap, penalties capped at 1e6
// (40_solver.metal:1689-1727)
__device__ __forceinline__ void dual_soft(const DeviceState& S, const StepParams& P, uint32_t gid) {
  SoftRec& sc = S.soft[gid];
  PoseView pv{S.posLin, S.posAng, S.posLin, S.posAng, false};
  V3 n, t1, t2, rAW;
  V3 C = soft_contact_C(sc, pv, 0u, S.initLin, S.initAng, P.alpha, n, t1, t2, rAW);
  V3 F = xyz(sc.penalty) * C + xyz(sc.lambda);
  F.x = fminf(F.x, 0.0f);
  float bnd = fabsf(F.x) * sc.normal.w;
  float fs = sqrtf(F.y * F.y + F.z * F.z);
  if (fs > bnd && fs > 0.0f) {
    F.y *= bnd / fs;
    F.z *= bnd / fs;
  }
  F.x = fmaxf(F.x, -sc.C0.w);
  sc.lambda = f4(F, sc.lambda.w);
  V3 pen = xyz(sc.penalty);
  if (F.x < 0.0f) pen.x = fminf(pen.x + P.betaLin * fabsf(C.x), kPenaltyMaxT);
  if (fs <= bnd) {
    pen.y = fminf(pen.y + P.betaLin * fabsf(C.y), kPenaltyMaxT);
    pen.z = fminf(pen.z + P.betaLin * fabsf(C.z), kPenaltyMaxT);
  }
  sc.penalty = f4(pen, sc.penalty.w);
}

// one item of the fused dual pass over [joints | springs | manifolds | soft],
// plus the next iteration's pose snapshot over bodies
__device__ __forceinline__ void dual_item(const DeviceState& S, const StepParams& P, uint32_t gid, uint32_t numPairs) {
  if (gid < P.numJoints) { dual_joint(S, P, gid); return; }
  uint32_t si = gid - P.numJoints;
  if (si < P.numSprings) { dual_spring(S, P, si); return; }
  uint32_t mi = si - P.numSprings;
  if (mi < numPairs * kMaxContacts) { dual_contact(S, P, mi / kMaxContacts, mi % kMaxContacts); return; }
  dual_soft(S, P, mi - numPairs * kMaxContacts);
}

__global__ void __launch_bounds__(256) k_dual(DeviceState S, StepParams P) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;
  uint32_t numPairs = S.counters[CTR_PAIRS];
  uint32_t total = P.numJoints + P.numSprings + numPairs * kMaxContacts + S.counters[CTR_SOFT];
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += stride) dual_item(S, P, gid, numPairs);
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < P.numBodies; b += stride) {
    S.posLinSnap[b] = S.posLin[b];
    S.posAngSnap[b] = S.posAng[b];
  }
}

void launch_dual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  uint64_t n = (uint64_t)P.numJoints + P.numSprings + (uint64_t)P.maxPairs * kMaxContacts + (P.numTris ? P.maxSoft : 0u);
  if (P.numBodies > n) n = P.numBodies;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_dual, dim3((uint32_t)blocks), dim3(256), 0, st, S, P);
}

// ---------------------------------------------------------------------------
// BDF1 velocity finalize (40_solver.metal:2108-2150)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_finalize(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numBodies) return;
  F4 vOld = S.velLin[gid];
  S.prevVelLin[gid] = vOld;
  F4 pl = S.posLin[gid];
  if (pl.w <= 0.0f) return;
  V3 v = ((xyz(pl) - xyz(S.initLin[gid])) + xyz(S.posLo[gid])) / P.dt;
  float s2 = dot(v, v);
  if (s2 > P.maxSpeed * P.maxSpeed) v *= P.maxSpeed * rsq(s2);
  float shw = S.shape[gid].w;
  if (shw < 0.0f && P.particleDamping > 0.0f) {
    V3 ballistic = xyz(vOld) + v3(0, 0, P.gravity * S.gravityScale[gid]) * P.dt;
    v = ballistic + (v - ballistic) / (1.0f + P.particleDamping * P.dt);
  }
  if (shw >= 0.0f && P.rigidLinearDamping > 0.0f) v *= expf(-P.rigidLinearDamping * P.dt);
  V3 omega = q_sub(S.posAng[gid], S.initAng[gid]) / P.dt;
  if (shw >= 0.0f && P.rigidAngularDamping > 0.0f) omega *= expf(-P.rigidAngularDamping * P.dt);
  S.velLin[gid] = f4(v, 0.0f);
  S.velAng[gid] = f4(omega, 0.0f);
}

void launch_finalize(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (P.numBodies == 0) return;
  hipLaunchKernelGGL(k_finalize, dim3((P.numBodies + 255) / 256), dim3(256), 0, st, S, P);
}

// ---------------------------------------------------------------------------
// residual: per dynamic body the block right-hand side at the current
// iterate (inertia + every stamp, no solve); block partials of |g|^2 and
// |m g_grav|^2 in residualPartial[2*block + {0,1}]
// ---------------------------------------------------------------------------
// guard: part of a residual-terminated frame's chunk, skipped once it has ended
__global__ void __launch_bounds__(256) k_residual(DeviceState S, StepParams P, uint32_t guard) {
  if (guard && S.counters[CTR_CONV_DONE] != 0u) return;
  __shared__ double red[2][256];
  uint32_t body = blockIdx.x * blockDim.x + threadIdx.x;
  double g2 = 0.0, w2 = 0.0;
  if (body < P.numBodies && S.posLin[body].w > 0.0f) {
    bool rigid = S.shape[body].w >= 0.0f;
    Acc acc;
    acc_zero(acc);
    PoseView pv{S.posLin, S.posAng, S.posLin, S.posAng, false};
    uint32_t ss = S.sAdjStart[body], sc = S.sAdjCount[body];
    uint32_t ds = S.dAdjStart[body], dc = S.dAdjCount[body];
    const V3 xs = xyz(S.posLin[body]), ls = xyz(S.posLo[body]);
    for (uint32_t k = 0; k < sc + dc; k++) {
      uint32_t entry = k < sc ? S.sAdjList[ss + k] : S.dAdjList[ds + (k - sc)];
      stamp_entry(S, pv, entry, ss + k, body, xs, ls, rigid, P.alpha, P.dt, acc);
    }
    F4 pl = S.posLin[body];
    float md = pl.w / (P.dt * P.dt);
    V3 g = acc.gl + ((xyz(pl) - xyz(S.inertLin[body])) + ls) * md;
    g2 = (double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z;
    if (rigid) {
      M3 inertia = m3_scale(world_inertia(S.posAng[body], xyz(S.props[body])), 1.0f / (P.dt * P.dt));
      V3 ga = acc.ga + m3_mul(inertia, q_sub(S.posAng[body], S.inertAng[body]));
      g2 += (double)ga.x * ga.x + (double)ga.y * ga.y + (double)ga.z * ga.z;
    }
    double mg = (double)pl.w * P.gravity * S.gravityScale[body];
    w2 = mg * mg;
  }
  red[0][threadIdx.x] = g2;
  red[1][threadIdx.x] = w2;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    S.residualPartial[2 * blockIdx.x] = red[0][0];
    S.residualPartial[2 * blockIdx.x + 1] = red[1][0];
  }
}

void launch_residual(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t blocks, bool guard) {
  hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, st, S, P, guard ? 1u : 0u);
}

// The residual check ending a chunk of `n` sweeps: r = sqrt(sum |g|^2 /
// sum |m g_grav|^2) from the block partials (f64); sets CTR_CONV_DONE when
// r <= tol (1) or the frame has run maxIt sweeps (2), and publishes the
// progress to the mapped host record.  A check after the frame has ended
// only advances the sequence number.
__global__ void __launch_bounds__(256) k_residual_check(DeviceState S, uint32_t blocks, float tol, uint32_t maxIt,
                                                        uint32_t n, ConvHost* host) {
  __shared__ double red[2][256];
  uint32_t* c = S.counters;
  const bool ended = c[CTR_CONV_DONE] != 0u;
  double g2 = 0.0, w2 = 0.0;
  if (!ended)
    for (uint32_t b = threadIdx.x; b < blocks; b += 256) {
      g2 += S.residualPartial[2 * b];
      w2 += S.residualPartial[2 * b + 1];
    }
  red[0][threadIdx.x] = g2;
  red[1][threadIdx.x] = w2;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const uint32_t seq = c[CTR_CONV_SEQ] + 1u;
  c[CTR_CONV_SEQ] = seq;
  if (!ended) {
    const double r = red[1][0] > 0.0 ? sqrt(red[0][0] / red[1][0]) : sqrt(red[0][0]);
    const uint32_t it = c[CTR_CONV_ITERS] + n;
    c[CTR_CONV_ITERS] = it;
    c[CTR_CONV_RES] = as_uint((float)r);
    c[CTR_CONV_DONE] = r <= (double)tol ? 1u : (it >= maxIt ? 2u : 0u);
  }
  host->done = c[CTR_CONV_DONE];
  host->iters = c[CTR_CONV_ITERS];
  host->residual = as_float(c[CTR_CONV_RES]);
  __hip_atomic_store(&host->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_residual_check(hipStream_t st, const DeviceState& S, uint32_t blocks, float tol, uint32_t maxIt,
                           uint32_t n, ConvHost* host) {
  hipLaunchKernelGGL(k_residual_check, dim3(1), dim3(256), 0, st, S, blocks, tol, maxIt, n, host);
}

// ---------------------------------------------------------------------------
// max constraint error (40_solver.metal:2182-2229)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_diag_error(DeviceState S, StepParams P, uint32_t* out) {
  uint32_t numPairs = S.counters[CTR_PAIRS];
  uint32_t total = P.numJoints + numPairs;
  for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += gridDim.x * blockDim.x) {
    float err = 0.0f;
    if (gid < P.numJoints) {
      const JointRec& j = S.joints[gid];
      if (j.header.z == 0 && (j.rA.w > 0.0f || j.rB.w > 0.0f)) {
        uint32_t a = j.header.x, b = j.header.y;
        V3 pA = a == kWorld ? xyz(j.rA) : xform(xyz(S.posLin[a]), S.posAng[a], xyz(j.rA));
        V3 pB = xform(xyz(S.posLin[b]), S.posAng[b], xyz(j.rB));
        err = length(pA - pB);
      }
    } else {
      const ManifoldRec& m = S.manifolds[gid - P.numJoints];
      uint32_t n = m.header.z, a = m.header.x, b = m.header.y;
      bool sA = (m.header.w & 2u) != 0, sB = (m.header.w & 4u) != 0;
      for (uint32_t i = 0; i < n; i++) {
        V3 xA = sA ? xyz(S.posLin[a]) + xyz(m.c[i].rA) : xform(xyz(S.posLin[a]), S.posAng[a], xyz(m.c[i].rA));
        V3 xB = sB ? xyz(S.posLin[b]) + xyz(m.c[i].rB) : xform(xyz(S.posLin[b]), S.posAng[b], xyz(m.c[i].rB));
        float pen = dot(xyz(m.basisN), xA - xB) + P.collisionMargin;
        err = fmaxf(err, fmaxf(0.0f, -pen));
      }
    }
    atomicMax(out, as_uint(err));
  }
}

void launch_diag_error(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t* out, uint32_t gridCap) {
  uint32_t blocks = (P.numJoints + P.maxPairs + 255) / 256;
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_diag_error, dim3(blocks), dim3(256), 0, st, S, P, out);
}

// ---------------------------------------------------------------------------
// persistent cooperative sweep: iterations x colors x dual in one launch,
// separated by grid barriers
// ---------------------------------------------------------------------------
template <int L, int MODE>
__global__ void __launch_bounds__(256) k_persistent(DeviceState S, StepParams P, uint32_t numColors,
                                                    uint32_t barrierBase) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;  // uniform across the grid: no barrier is entered
  const uint32_t groupsPerBlock = 256 / L;
  const uint32_t lane = threadIdx.x & (L - 1);
  const uint32_t gslot = xcd_block(blockIdx.x, gridDim.x) * groupsPerBlock + threadIdx.x / L;
  const uint32_t ngroups = gridDim.x * groupsPerBlock;
  uint32_t* bar = &S.counters[CTR_GBAR];
  uint32_t* fail = &S.counters[CTR_GBAR_FAIL];
  uint32_t epoch = barrierBase;
  const uint32_t colorsUsed = numColors ? numColors : S.colorStart[kMaxColors + 1];
  const uint32_t numPairs = S.counters[CTR_PAIRS];
  const uint32_t dualTotal = P.numJoints + P.numSprings + numPairs * kMaxContacts + S.counters[CTR_SOFT];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
  for (uint32_t it = 0; it < P.iterations; it++) {
    for (uint32_t c = 0; c < colorsUsed; c++) {
      uint32_t s = S.colorStart[c], e = S.colorStart[c + 1];
      uint32_t count = e - s;
      uint32_t rounds = (count + ngroups - 1) / ngroups;
      for (uint32_t r = 0; r < rounds; r++) {
        uint32_t slot = r * ngroups + gslot;
        bool active = slot < count;
        primal_group<L, MODE>(S, P, active,
                              numColors ? (active ? S.colorWork[s + slot] : U4{0u, 0u, 0u, 0u})
                                        : body_work(S, active, active ? S.colorList[s + slot] : 0u),
                              lane, c);
      }
      epoch += gridDim.x;
      grid_barrier(bar, fail, epoch);
    }
    for (uint32_t g = tid; g < dualTotal; g += nthreads) dual_item(S, P, g, numPairs);
    for (uint32_t b = tid; b < P.numBodies; b += nthreads) {
      S.posLinSnap[b] = S.posLin[b];
      S.posAngSnap[b] = S.posAng[b];
    }
    epoch += gridDim.x;
    grid_barrier(bar, fail, epoch);
  }
}

template <int MODE>
static const void* persistent_fn_t(int lanes) {
  switch (lanes) {
    case 4: return (const void*)k_persistent<4, MODE>;
    case 8: return (const void*)k_persistent<8, MODE>;
    case 16: return (const void*)k_persistent<16, MODE>;
    default: return (const void*)k_persistent<32, MODE>;
  }
}
static const void* persistent_fn(int lanes, int mode) {
  if (mode == kModeTet) return persistent_fn_t<kModeTet>(lanes);
  if (mode == kModeParticle) return persistent_fn_t<kModeParticle>(lanes);
  return persistent_fn_t<kModeGeneral>(lanes);
}

int persistent_max_blocks(int lanes, int mode) {
  int dev = 0, cus = 0, per = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, persistent_fn(lanes, mode), 256, 0);
  if (e != hipSuccess || per <= 0 || cus <= 0) return 0;
  return per * cus;
}

hipError_t launch_persistent_solve(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t numColors,
                                   int lanes, int mode, uint32_t gridBlocks, uint32_t barrierBase) {
  DeviceState s = S;
  StepParams p = P;
  uint32_t nc = numColors, bb = barrierBase;
  void* args[] = {&s, &p, &nc, &bb};
  return hipLaunchCooperativeKernel(persistent_fn(lanes, mode), dim3(gridBlocks), dim3(256), args, 0, st);
}

}  // namespace avbd
