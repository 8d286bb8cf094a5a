// cloth.hip -- cloth-surface contact pipeline for gfx950: the element grid,
// the Planar-DAT vertex-triangle / edge-edge pair stream, OGC element-contact
// emission (V-T, E-E and rigid feature vs triangle) with the soft
// persistence map, and the 1-ring velocity viscosity.
//
// Follows 25_cloth.metal:1-333, 990-1163, 27_planar_dat.metal:1-510 and
// 40_solver.metal:2154-2177.  The reference appends pairs and contacts with
// atomics, so its record order depends on the schedule; here every stream is
// built count / scan / emit (pairs additionally sorted per producer), so
// record order -- and with it every downstream sum -- is a function of the
// inputs only.
#include "kernels.hpp"
#include "grid.hpp"

namespace avbd {

namespace {

constexpr float kFltMax = 3.402823466e38f;
constexpr uint32_t kRtMaxPerTri = 4u;
constexpr float kEeInteriorEps = 0.02f;
constexpr float kDatGeometryEps = 1.0e-8f;
constexpr uint32_t kDatLdsCap = 512u;  // per-producer sorted-emission buffer (LDS)

// Ericson, Real-Time Collision Detection 5.1.5 (25_cloth.metal:26-61)
__device__ __forceinline__ V3 closest_pt_triangle(V3 p, V3 a, V3 b, V3 c, V3& bary) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { bary = v3(1, 0, 0); return a; }
  V3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { bary = v3(0, 1, 0); return b; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / fmaxf(d1 - d3, 1e-12f);
    bary = v3(1 - v, v, 0);
    return a + ab * v;
  }
  V3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { bary = v3(0, 0, 1); return c; }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / fmaxf(d2 - d6, 1e-12f);
    bary = v3(1 - w, 0, w);
    return a + ac * w;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / fmaxf((d4 - d3) + (d5 - d6), 1e-12f);
    bary = v3(0, 1 - w, w);
    return b + (c - b) * w;
  }
  float denom = va + vb + vc;
  if (fabsf(denom) < 1e-20f) { bary = v3(1, 0, 0); return a; }
  float v = vb / denom, w = vc / denom;
  bary = v3(1 - v - w, v, w);
  return a + ab * v + ac * w;
}

// closest points of two segments, parameters s on p, t on q (25_cloth.metal:632-648)
__device__ __forceinline__ void closest_seg_seg(V3 p0, V3 p1, V3 q0, V3 q1, float& s, float& t) {
  V3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  s = 0.0f;
  t = 0.0f;
  if (a <= 1e-12f && e <= 1e-12f) return;
  if (a <= 1e-12f) { t = clampf(f / e, 0.0f, 1.0f); return; }
  float c = dot(d1, r);
  if (e <= 1e-12f) { s = clampf(-c / a, 0.0f, 1.0f); return; }
  float b = dot(d1, d2);
  float denom = a * e - b * b;
  if (fabsf(denom) > 1e-12f) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
  t = (b * s + f) / e;
  if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
  else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
}

__device__ __forceinline__ bool nbr_contains(const uint32_t* list, uint32_t s, uint32_t e, uint32_t x) {
  uint32_t lo = s, hi = e;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (list[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < e && list[lo] == x;
}

__device__ __forceinline__ V3 P3(const DeviceState& S, uint32_t b) { return xyz(S.posLin[b]); }

// Element AABB padded by the pair-query radius (25_cloth.metal:147-181,
// Planar-DAT branch): triangles first, then edges.
__device__ __forceinline__ void elem_bounds(const DeviceState& S, const StepParams& P, uint32_t gid, V3& lo, V3& hi) {
  V3 pad = v3s(P.datQueryRadius);
  if (gid < P.numTris) {
    U4 t = S.tris[gid];
    V3 a = P3(S, t.x), b = P3(S, t.y), c = P3(S, t.z);
    lo = vmin(vmin(a, b), c) - pad;
    hi = vmax(vmax(a, b), c) + pad;
  } else {
    U2 e = S.edges[gid - P.numTris];
    V3 a = P3(S, e.x), b = P3(S, e.y);
    lo = vmin(a, b) - pad;
    hi = vmax(a, b) + pad;
  }
}

struct Cell3 {
  int x, y, z;
};
__device__ __forceinline__ Cell3 cell_of(V3 p, float cs) {
  Cell3 c;
  cell_coord(p, cs, c.x, c.y, c.z);
  return c;
}
// Cells c0 .. c0 + span of an AABB, the span clamped to 3 per axis (a cover
// of at most 4 cells); computed in 64 bits so saturated coordinates of a
// diverged state cannot wrap into unbounded loops.
struct CellRange {
  Cell3 c0;
  int sx, sy, sz;
  bool over;  // the unclamped cover exceeded 4 cells on an axis
};
__device__ __forceinline__ int clamp_span(long long d) { return d < 0 ? 0 : (d > 3 ? 3 : (int)d); }
__device__ __forceinline__ CellRange range_of(V3 lo, V3 hi, float cs) {
  CellRange r;
  r.c0 = cell_of(lo, cs);
  Cell3 raw = cell_of(hi, cs);
  long long dx = (long long)raw.x - r.c0.x, dy = (long long)raw.y - r.c0.y, dz = (long long)raw.z - r.c0.z;
  r.over = dx > 3 || dy > 3 || dz > 3;
  r.sx = clamp_span(dx);
  r.sy = clamp_span(dy);
  r.sz = clamp_span(dz);
  return r;
}
__device__ __forceinline__ int cell_add(int c, int i) { return (int)((uint32_t)c + (uint32_t)i); }
__device__ __forceinline__ bool axis_in(int c, int c0, int s) {
  return (long long)c >= c0 && (long long)c <= (long long)c0 + s;
}
__device__ __forceinline__ bool in_range(const CellRange& r, Cell3 c) {
  return axis_in(c.x, r.c0.x, r.sx) && axis_in(c.y, r.c0.y, r.sy) && axis_in(c.z, r.c0.z, r.sz);
}

}  // namespace

// ---------------------------------------------------------------------------
// element grid: every triangle / edge registers in each distinct hash bucket
// of its padded AABB cover, at most 4 cells per axis (25_cloth.metal:182-267)
// ---------------------------------------------------------------------------
namespace {

// Bucket counters aggregated per workgroup: an LDS table keyed by bucket slot
// gathers the block's (element, cell) registrations, then one device atomic
// per distinct slot adds the block's count (count pass) or reserves the
// block's run of the bucket (scatter pass).  A block of consecutive elements
// of a sheet touches a few hundred buckets, so the device-scope atomics drop
// roughly tenfold.  A slot that finds no table entry within kBinProbes
// probes goes to the device counter directly (count pass) or turns the whole
// block to direct per-entry cursors (scatter pass), so a full table costs
// speed only.  Entries land in a bucket in no particular order, as before;
// every consumer of the buckets is order independent.
constexpr uint32_t kBinTable = 2048u, kBinProbes = 32u, kBinEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t bin_entry(uint32_t* key, uint32_t slot, bool insert) {
  uint32_t i = (slot * 2654435761u) >> 21;
  for (uint32_t p = 0; p < kBinProbes; p++, i = (i + 1u) & (kBinTable - 1u)) {
    uint32_t k = __hip_atomic_load(&key[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == slot) return i;
    if (k != kBinEmpty) continue;
    if (!insert) return kBinEmpty;
    uint32_t old = atomicCAS(&key[i], kBinEmpty, slot);
    if (old == kBinEmpty || old == slot) return i;
  }
  return kBinEmpty;
}

// f(slot) once per distinct hash bucket of the cover (a cover of up to 64
// cells can hash two cells to one bucket; the element registers there once)
template <typename F>
__device__ __forceinline__ void for_each_bucket(const StepParams& P, const CellRange& r, bool tri, F&& f) {
  const Cell3 c0 = r.c0;
  const int nx = r.sx + 1, ny = r.sy + 1, nz = r.sz + 1;
  const int n = nx * ny * nz;  // <= 64
  // `seen` holds one bit per bucket (mod 64) already registered: a cell whose
  // bit is clear cannot repeat a bucket, so only the rare candidates of a
  // repeat rescan the cover's earlier cells
  uint64_t seen = 0ull;
  for (int i = 0; i < n; i++) {
    int x = cell_add(c0.x, i % nx), y = cell_add(c0.y, (i / nx) % ny), z = cell_add(c0.z, i / (nx * ny));
    uint32_t h = cell_hash(x, y, z, 0u, P.elemHashSize);
    const uint64_t bit = 1ull << (h & 63u);
    bool dup = false;
    if ((seen & bit) != 0ull)
      for (int j = 0; j < i && !dup; j++) {
        int xj = cell_add(c0.x, j % nx), yj = cell_add(c0.y, (j / nx) % ny), zj = cell_add(c0.z, j / (nx * ny));
        dup = cell_hash(xj, yj, zj, 0u, P.elemHashSize) == h;
      }
    seen |= bit;
    // counts and cursors are (triangles, edges) per bucket, interleaved, so
    // one scan lays out every bucket as its triangles, then its edges
    if (!dup) f(2u * h + (tri ? 0u : 1u));
  }
}

}  // namespace

template <bool SCATTER>
__global__ void __launch_bounds__(256) k_el_bin(DeviceState S, StepParams P) {
  __shared__ uint32_t key[kBinTable], cnt[kBinTable];
  __shared__ uint32_t spill;
  for (uint32_t i = threadIdx.x; i < kBinTable; i += blockDim.x) {
    key[i] = kBinEmpty;
    cnt[i] = 0u;
  }
  if (threadIdx.x == 0) spill = 0u;
  __syncthreads();
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = gid < P.numTris + P.numEdges;
  const bool tri = gid < P.numTris;
  CellRange r{};
  if (live) {
    V3 lo, hi;
    elem_bounds(S, P, gid, lo, hi);
    if (!SCATTER) {
      // lo.w carries the element's surface group when all its vertices share
      // one (else 0): the queries test the same-component filter from it
      uint32_t ge;
      if (tri) {
        U4 t = S.tris[gid];
        ge = S.clothGroup[t.x];
        if (S.clothGroup[t.y] != ge || S.clothGroup[t.z] != ge) ge = 0u;
      } else {
        U2 e = S.edges[gid - P.numTris];
        ge = S.clothGroup[e.x];
        if (S.clothGroup[e.y] != ge) ge = 0u;
      }
      S.elemBox[2 * gid] = f4(lo, as_float(ge));
      S.elemBox[2 * gid + 1] = f4(hi, 0.0f);
    }
    if (!finite3(lo) || !finite3(hi)) {
      if (!SCATTER) {
        atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
        atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
      }
      live = false;
    } else {
      r = range_of(lo, hi, P.elemCellSize);
      if (!SCATTER && r.over) atomicAdd(&S.counters[CTR_DAT_OVERFLOW], 1u);
    }
  }
  if (live)
    for_each_bucket(P, r, tri, [&](uint32_t slot) {
      uint32_t i = bin_entry(key, slot, true);
      if (i != kBinEmpty) atomicAdd(&cnt[i], 1u);
      else if (SCATTER) spill = 1u;
      else atomicAdd(&S.elemCellCount[slot], 1u);
    });
  __syncthreads();
  if (!SCATTER) {
    for (uint32_t i = threadIdx.x; i < kBinTable; i += blockDim.x)
      if (key[i] != kBinEmpty) atomicAdd(&S.elemCellCount[key[i]], cnt[i]);
    return;
  }
  const bool direct = spill != 0u;
  if (!direct)
    for (uint32_t i = threadIdx.x; i < kBinTable; i += blockDim.x)
      if (key[i] != kBinEmpty) cnt[i] = atomicAdd(&S.elemCellCursor[key[i]], cnt[i]);
  __syncthreads();
  if (!live) return;
  const uint32_t entry = tri ? gid : (kElemEdgeBit | (gid - P.numTris));
  for_each_bucket(P, r, tri, [&](uint32_t slot) {
    uint32_t pos = direct ? atomicAdd(&S.elemCellCursor[slot], 1u) : atomicAdd(&cnt[bin_entry(key, slot, false)], 1u);
    S.elemCells[pos] = entry;
  });
}

__global__ void __launch_bounds__(256) k_copy_u32(uint32_t* dst, const uint32_t* src, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Planar-DAT pair stream (27_planar_dat.metal:35-229): producer p is surface
// particle p (V-T query over its cell's bucket) or, past the particles, edge
// p - numParticles (E-E over the buckets of its tight AABB, owned by A < B
// and by the minimum cell of tight(A) x padded(B)).
// ---------------------------------------------------------------------------
namespace {

// Early rejection by the element's padded box (its tight box grown by the
// query radius, stored by k_el_bin): a partner within the radius lies inside
// it.  The test box is grown by a further kBoxSlack so rounding of the stored
// box can never reject a pair the exact test accepts.
constexpr float kBoxSlack = 1.0e-5f;
__device__ __forceinline__ bool outside_box(V3 p, F4 lo, F4 hi, float slack) {
  return p.x < lo.x - slack || p.y < lo.y - slack || p.z < lo.z - slack || p.x > hi.x + slack ||
         p.y > hi.y + slack || p.z > hi.z + slack;
}
__device__ __forceinline__ bool boxes_apart(V3 alo, V3 ahi, F4 lo, F4 hi, float slack) {
  return ahi.x < lo.x - slack || ahi.y < lo.y - slack || ahi.z < lo.z - slack || alo.x > hi.x + slack ||
         alo.y > hi.y + slack || alo.z > hi.z + slack;
}

// Each test is split in two: the box filter (the candidate's stored padded
// box and surface group, one 32-byte read) and the rest (connectivity, the
// owner cell, exact distance, 1-ring topology).  The accept predicate is
// their conjunction; the count pass queues the box-filter survivors of a
// wave's producers and runs the rest on full waves.
//
// `noSelf`: the producer's surface group when that surface does not
// self-collide (else 0); a candidate of that group fails the same-component
// filter, tested first from the group stored with the element's box.
__device__ __forceinline__ bool vt_box(const DeviceState& S, const StepParams& P, V3 p, uint32_t entry,
                                       uint32_t noSelf) {
  if ((entry & kElemEdgeBit) != 0u || entry >= P.numTris) return false;
  const F4 blo = S.elemBox[2 * entry], bhi = S.elemBox[2 * entry + 1];
  if (noSelf != 0u && as_uint(blo.w) == noSelf) return false;
  return !outside_box(p, blo, bhi, kBoxSlack);
}

__device__ __forceinline__ bool vt_rest(const DeviceState& S, const StepParams& P, uint32_t v, V3 p, Cell3 cell,
                                        uint32_t ns, uint32_t ne, uint32_t entry, float q2) {
  U4 t = S.tris[entry];
  if (t.x == v || t.y == v || t.z == v) return false;
  const F4 blo = S.elemBox[2 * entry], bhi = S.elemBox[2 * entry + 1];
  if (!in_range(range_of(xyz(blo), xyz(bhi), P.elemCellSize), cell)) return false;
  V3 bary;
  V3 q = closest_pt_triangle(p, P3(S, t.x), P3(S, t.y), P3(S, t.z), bary);
  V3 d = p - q;
  if (dot(d, d) > q2) return false;
  // topology filters last: most candidates of a bucket are out of reach, and
  // the 1-ring searches are the longest dependent-load chains
  uint32_t g = S.clothGroup[v];
  bool same = g != 0u && S.clothGroup[t.x] == g && S.clothGroup[t.y] == g && S.clothGroup[t.z] == g;
  if (same && S.selfCollision[v] == 0u) return false;
  return !(same && (nbr_contains(S.nbrList, ns, ne, t.x) || nbr_contains(S.nbrList, ns, ne, t.y) ||
                    nbr_contains(S.nbrList, ns, ne, t.z)));
}

__device__ __forceinline__ bool vt_accept(const DeviceState& S, const StepParams& P, uint32_t v, V3 p, Cell3 cell,
                                          uint32_t ns, uint32_t ne, uint32_t entry, float q2, uint32_t noSelf) {
  return vt_box(S, P, p, entry, noSelf) && vt_rest(S, P, v, p, cell, ns, ne, entry, q2);
}

__device__ __forceinline__ bool ee_box(const DeviceState& S, const StepParams& P, uint32_t gid, V3 a0, V3 a1,
                                       uint32_t entry, uint32_t noSelf) {
  if ((entry & kElemEdgeBit) == 0u) return false;
  uint32_t other = entry & ~kElemEdgeBit;
  if (other <= gid || other >= P.numEdges) return false;
  const F4 blo = S.elemBox[2 * (P.numTris + other)], bhi = S.elemBox[2 * (P.numTris + other) + 1];
  if (noSelf != 0u && as_uint(blo.w) == noSelf) return false;
  return !boxes_apart(vmin(a0, a1), vmax(a0, a1), blo, bhi, kBoxSlack);
}

__device__ __forceinline__ bool ee_rest(const DeviceState& S, const StepParams& P, U2 a, V3 a0, V3 a1,
                                        const CellRange& ar, uint32_t ns0, uint32_t ne0, uint32_t ns1, uint32_t ne1,
                                        Cell3 cell, uint32_t other, float q2) {
  U2 b = S.edges[other];
  if (b.x == a.x || b.x == a.y || b.y == a.x || b.y == a.y) return false;
  const F4 blo = S.elemBox[2 * (P.numTris + other)], bhi = S.elemBox[2 * (P.numTris + other) + 1];
  CellRange br = range_of(xyz(blo), xyz(bhi), P.elemCellSize);
  // owned by the minimum cell of tight(A) x padded(B)
  Cell3 o0{max(ar.c0.x, br.c0.x), max(ar.c0.y, br.c0.y), max(ar.c0.z, br.c0.z)};
  if (!in_range(ar, o0) || !in_range(br, o0)) return false;
  if (cell.x != o0.x || cell.y != o0.y || cell.z != o0.z) return false;
  float s, t;
  V3 b0 = P3(S, b.x), b1 = P3(S, b.y);
  closest_seg_seg(a0, a1, b0, b1, s, t);
  V3 ca = a0 + (a1 - a0) * s, cb = b0 + (b1 - b0) * t;
  V3 d = ca - cb;
  if (dot(d, d) > q2) return false;
  // topology filters last (see vt_rest)
  uint32_t g = S.clothGroup[a.x];
  bool same = g != 0u && S.clothGroup[a.y] == g && S.clothGroup[b.x] == g && S.clothGroup[b.y] == g;
  if (same && S.selfCollision[a.x] == 0u) return false;
  return !(same && (nbr_contains(S.nbrList, ns0, ne0, b.x) || nbr_contains(S.nbrList, ns0, ne0, b.y) ||
                    nbr_contains(S.nbrList, ns1, ne1, b.x) || nbr_contains(S.nbrList, ns1, ne1, b.y)));
}

__device__ __forceinline__ bool ee_accept(const DeviceState& S, const StepParams& P, uint32_t gid, U2 a, V3 a0, V3 a1,
                                          const CellRange& ar, uint32_t ns0, uint32_t ne0, uint32_t ns1, uint32_t ne1,
                                          Cell3 cell, uint32_t entry, float q2, uint32_t noSelf) {
  return ee_box(S, P, gid, a0, a1, entry, noSelf) &&
         ee_rest(S, P, a, a0, a1, ar, ns0, ne0, ns1, ne1, cell, entry & ~kElemEdgeBit, q2);
}

// Visits producer `prod`'s candidates in chunks of `width` positions starting
// at `lane`; f(accepted, other) runs once per chunk on every lane (uniform
// trip counts, so a wave may ballot inside f).  Counters only when `diag`.
template <typename F>
__device__ void dat_visit(const DeviceState& S, const StepParams& P, uint32_t prod, uint32_t lane, uint32_t width,
                          bool diag, F&& f) {
  const float q2 = P.datQueryRadius * P.datQueryRadius;
  if (prod < P.numParticles) {
    uint32_t v = S.particleIdx[prod];
    V3 p = P3(S, v);
    if (!finite3(p)) {
      if (diag && lane == 0) {
        atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
        atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
      }
      return;
    }
    Cell3 cell = cell_of(p, P.elemCellSize);
    uint32_t h = cell_hash(cell.x, cell.y, cell.z, 0u, P.elemHashSize);
    uint32_t s = S.elemCellStart[2u * h], e = S.elemCellStart[2u * h + 1u];  // the bucket's triangles
    uint32_t ns = S.nbrStart[v], ne = ns + S.nbrCount[v];
    const uint32_t noSelf = S.selfCollision[v] == 0u ? S.clothGroup[v] : 0u;
    for (uint32_t base = s; base < e; base += width) {
      uint32_t k = base + lane;
      uint32_t entry = k < e ? S.elemCells[k] : 0u;
      bool acc = k < e && vt_accept(S, P, v, p, cell, ns, ne, entry, q2, noSelf);
      f(acc, entry);
    }
    return;
  }
  uint32_t gid = prod - P.numParticles;
  U2 a = S.edges[gid];
  V3 a0 = P3(S, a.x), a1 = P3(S, a.y);
  if (!finite3(a0) || !finite3(a1)) {
    if (diag && lane == 0) {
      atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
      atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
    }
    return;
  }
  CellRange ar = range_of(vmin(a0, a1), vmax(a0, a1), P.elemCellSize);
  if (diag && lane == 0 && ar.over) atomicAdd(&S.counters[CTR_DAT_OVERFLOW], 1u);
  uint32_t ns0 = S.nbrStart[a.x], ne0 = ns0 + S.nbrCount[a.x];
  uint32_t ns1 = S.nbrStart[a.y], ne1 = ns1 + S.nbrCount[a.y];
  const uint32_t ga = S.clothGroup[a.x];
  const uint32_t noSelf = (S.selfCollision[a.x] == 0u && S.clothGroup[a.y] == ga) ? ga : 0u;
  for (int iz = 0; iz <= ar.sz; iz++)
    for (int iy = 0; iy <= ar.sy; iy++)
      for (int ix = 0; ix <= ar.sx; ix++) {
        int x = cell_add(ar.c0.x, ix), y = cell_add(ar.c0.y, iy), z = cell_add(ar.c0.z, iz);
        Cell3 cell{x, y, z};
        uint32_t h = cell_hash(x, y, z, 0u, P.elemHashSize);
        uint32_t s = S.elemCellStart[2u * h + 1u], e = S.elemCellStart[2u * h + 2u];  // the bucket's edges
        for (uint32_t base = s; base < e; base += width) {
          uint32_t k = base + lane;
          uint32_t entry = k < e ? S.elemCells[k] : 0u;
          bool acc = k < e && ee_accept(S, P, gid, a, a0, a1, ar, ns0, ne0, ns1, ne1, cell, entry, q2, noSelf);
          f(acc, entry & ~kElemEdgeBit);
        }
      }
}

__device__ __forceinline__ U4 dat_record(const StepParams& P, uint32_t prod, uint32_t other) {
  return prod < P.numParticles ? U4{kDatVT, 0u, other, 0u} : U4{kDatEE, prod - P.numParticles, other, 0u};
}

}  // namespace

// Emit pass: one wave (64 lanes) per producer with pairs, four producers
// per block; the accepted partners are gathered in LDS and written ranked by
// partner index at datStart[prod] (k_dat_count below is the count pass).
__global__ void __launch_bounds__(256) k_dat_pairs(DeviceState S, StepParams P) {
  __shared__ uint32_t buf[4][kDatLdsCap];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t prod = blockIdx.x * 4u + wv;
  const uint32_t nProd = P.numParticles + P.numEdges;
  const bool valid = prod < nProd;
  uint32_t start = 0, want = 0;
  if (valid) {
    start = S.datStart[prod];
    want = S.datStart[prod + 1] - start;
  }
  const bool big = want > kDatLdsCap;
  uint32_t n = 0;
  if (valid && want != 0u && !big) {
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    dat_visit(S, P, prod, lane, 64u, false, [&](bool acc, uint32_t other) {
      uint64_t m = __ballot(acc);
      if (acc) {
        uint32_t r = n + (uint32_t)__popcll(m & below);
        if (r < kDatLdsCap) buf[wv][r] = other;
      }
      n += (uint32_t)__popcll(m);
    });
  }
  __syncthreads();
  if (!valid || want == 0u) return;
  const uint32_t wo = prod < P.numParticles ? S.particleIdx[prod] : 0u;
  if (!big) {
    uint32_t cnt = n < want ? n : want;
    for (uint32_t i = lane; i < cnt; i += 64u) {
      uint32_t x = buf[wv][i], rank = 0;
      for (uint32_t j = 0; j < cnt; j++) rank += buf[wv][j] < x ? 1u : 0u;
      uint32_t slot = start + rank;
      if (slot < P.maxDatPairs) {
        U4 r = dat_record(P, prod, x);
        if (r.x == kDatVT) r.y = wo;
        S.datPairs[slot] = r;
      }
    }
    return;
  }
  // more partners than the LDS buffer holds: one lane inserts in order
  if (lane != 0) return;
  uint32_t cnt = 0;
  dat_visit(S, P, prod, 0u, 1u, false, [&](bool acc, uint32_t other) {
    if (!acc || cnt >= want) return;
    uint32_t j = cnt;
    while (j > 0 && start + j - 1 < P.maxDatPairs && S.datPairs[start + j - 1].z > other) {
      if (start + j < P.maxDatPairs) S.datPairs[start + j] = S.datPairs[start + j - 1];
      j--;
    }
    if (start + j < P.maxDatPairs) {
      U4 r = dat_record(P, prod, other);
      if (r.x == kDatVT) r.y = wo;
      S.datPairs[start + j] = r;
    }
    cnt++;
  });
}

// Count pass: 16 lanes per producer, four producers per wave.  The groups
// walk their producers' buckets 16 candidates at a time through the box
// filter only; the survivors go to the wave's LDS queue tagged with their
// producer (and, for E-E, the cell whose bucket held them), and whenever
// the queue holds 64 of them the remaining tests run on the full wave.  The
// accepted count is the same as the emit pass's walk (both evaluate the same
// predicate on the same candidates); the queue only keeps the long tests off
// lanes the box filter already emptied.
namespace {

constexpr uint32_t kDatQueue = 128u;  // < 64 held + up to 64 added per step

struct DatProducer {
  uint32_t kind, id, ax, ay, ns0, ne0, ns1, ne1;  // kind 0: V-T (id = vertex), 1: E-E (id = edge)
  float p0[3], p1[3];                             // vertex / edge ends
  int c0[3], span[3];                             // the V-T cell, or the edge's tight cover
};

__device__ __forceinline__ Cell3 cover_cell(const CellRange& r, uint32_t ci) {
  const uint32_t nx = (uint32_t)r.sx + 1u, ny = (uint32_t)r.sy + 1u;
  return Cell3{cell_add(r.c0.x, (int)(ci % nx)), cell_add(r.c0.y, (int)((ci / nx) % ny)),
               cell_add(r.c0.z, (int)(ci / (nx * ny)))};
}

}  // namespace

// Six waves per SIMD (80 VGPRs, four spilled) measured 1 % faster on the
// 256^2 cloth step than the compiler's choice of five (85 VGPRs); eight
// spill 36 registers.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_dat_count(DeviceState S, StepParams P) {
  __shared__ DatProducer ctx[16];
  __shared__ uint32_t qEntry[4][kDatQueue], qTag[4][kDatQueue];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, sub = threadIdx.x & 15u, grp = lane >> 4;
  const uint32_t slot = threadIdx.x >> 4;
  const uint32_t prod = blockIdx.x * 16u + slot;
  const uint32_t nProd = P.numParticles + P.numEdges;
  const float q2 = P.datQueryRadius * P.datQueryRadius;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // producer set-up: every lane of a group computes the same values
  const bool vt = prod < P.numParticles;
  bool live = prod < nProd;
  V3 p0 = v3(0.0f, 0.0f, 0.0f), p1 = p0;
  uint32_t gid = 0, noSelf = 0, nCells = 0;
  CellRange ar{};
  if (live && vt) {
    const uint32_t v = S.particleIdx[prod];
    p0 = P3(S, v);
    if (!finite3(p0)) {
      if (sub == 0) {
        atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
        atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
      }
      live = false;
    } else {
      ar.c0 = cell_of(p0, P.elemCellSize);
      nCells = 1;
      noSelf = S.selfCollision[v] == 0u ? S.clothGroup[v] : 0u;
      if (sub == 0) {
        DatProducer& c = ctx[slot];
        const uint32_t ns = S.nbrStart[v];
        c.kind = 0u; c.id = v; c.ns0 = ns; c.ne0 = ns + S.nbrCount[v];
        c.p0[0] = p0.x; c.p0[1] = p0.y; c.p0[2] = p0.z;
        c.c0[0] = ar.c0.x; c.c0[1] = ar.c0.y; c.c0[2] = ar.c0.z;
      }
    }
  } else if (live) {
    gid = prod - P.numParticles;
    const U2 a = S.edges[gid];
    p0 = P3(S, a.x);
    p1 = P3(S, a.y);
    if (!finite3(p0) || !finite3(p1)) {
      if (sub == 0) {
        atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
        atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
      }
      live = false;
    } else {
      ar = range_of(vmin(p0, p1), vmax(p0, p1), P.elemCellSize);
      if (sub == 0 && ar.over) atomicAdd(&S.counters[CTR_DAT_OVERFLOW], 1u);
      nCells = (uint32_t)((ar.sx + 1) * (ar.sy + 1) * (ar.sz + 1));
      const uint32_t ga = S.clothGroup[a.x];
      noSelf = (S.selfCollision[a.x] == 0u && S.clothGroup[a.y] == ga) ? ga : 0u;
      if (sub == 0) {
        DatProducer& c = ctx[slot];
        const uint32_t ns0 = S.nbrStart[a.x], ns1 = S.nbrStart[a.y];
        c.kind = 1u; c.id = gid; c.ax = a.x; c.ay = a.y;
        c.ns0 = ns0; c.ne0 = ns0 + S.nbrCount[a.x]; c.ns1 = ns1; c.ne1 = ns1 + S.nbrCount[a.y];
        c.p0[0] = p0.x; c.p0[1] = p0.y; c.p0[2] = p0.z;
        c.p1[0] = p1.x; c.p1[1] = p1.y; c.p1[2] = p1.z;
        c.c0[0] = ar.c0.x; c.c0[1] = ar.c0.y; c.c0[2] = ar.c0.z;
        c.span[0] = ar.sx; c.span[1] = ar.sy; c.span[2] = ar.sz;
      }
    }
  }
  __syncthreads();

  // a V-T producer reads its cell's triangles, an E-E producer each cover
  // cell's edges (the buckets lay out triangles, then edges)
  auto bucket = [&](uint32_t ci, uint32_t& bs, uint32_t& be) {
    const Cell3 c = cover_cell(ar, ci);
    const uint32_t h = cell_hash(c.x, c.y, c.z, 0u, P.elemHashSize);
    const uint32_t o = vt ? 0u : 1u;
    bs = S.elemCellStart[2u * h + o];
    be = S.elemCellStart[2u * h + o + 1u];
  };
  // the remaining tests on queue items [0, n), then the queue's tail moves
  // to its head; `count` gathers the accepted items of this lane's producer
  uint32_t qn = 0, count = 0;
  auto drain = [&](uint32_t n) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool has = lane < n;
    const uint32_t entry = has ? qEntry[wv][lane] : 0u, tag = has ? qTag[wv][lane] : 0u;
    const uint32_t g = tag & 3u;
    bool acc = false;
    if (has) {
      const DatProducer& c = ctx[wv * 4u + g];
      const V3 x0 = v3(c.p0[0], c.p0[1], c.p0[2]);
      CellRange cr;
      cr.c0 = Cell3{c.c0[0], c.c0[1], c.c0[2]};
      if (c.kind == 0u) {
        acc = vt_rest(S, P, c.id, x0, cr.c0, c.ns0, c.ne0, entry, q2);
      } else {
        cr.sx = c.span[0];
        cr.sy = c.span[1];
        cr.sz = c.span[2];
        cr.over = false;
        const V3 x1 = v3(c.p1[0], c.p1[1], c.p1[2]);
        acc = ee_rest(S, P, U2{c.ax, c.ay}, x0, x1, cr, c.ns0, c.ne0, c.ns1, c.ne1, cover_cell(cr, tag >> 2),
                      entry & ~kElemEdgeBit, q2);
      }
    }
    for (uint32_t gg = 0; gg < 4u; gg++) {
      const uint64_t m = __ballot(acc && g == gg);
      if (gg == grp) count += (uint32_t)__popcll(m);
    }
    if (qn > n) {
      const bool mv = lane < qn - n;
      const uint32_t me = mv ? qEntry[wv][n + lane] : 0u, mt = mv ? qTag[wv][n + lane] : 0u;
      __builtin_amdgcn_wave_barrier();
      if (mv) {
        qEntry[wv][lane] = me;
        qTag[wv][lane] = mt;
      }
    }
    qn -= n;
  };

  uint32_t ci = 0, k0 = 0, ke = 0;
  bool more = live;
  if (more) bucket(0u, k0, ke);
  while (true) {
    while (more && k0 >= ke) {
      if (++ci >= nCells) more = false;
      else bucket(ci, k0, ke);
    }
    if (__ballot(more) == 0ull) break;
    const uint32_t k = k0 + sub;
    uint32_t entry = 0u;
    bool surv = false;
    if (more && k < ke) {
      entry = S.elemCells[k];
      surv = vt ? vt_box(S, P, p0, entry, noSelf) : ee_box(S, P, gid, p0, p1, entry, noSelf);
    }
    k0 += 16u;
    const uint64_t m = __ballot(surv);
    if (surv) {
      const uint32_t pos = qn + (uint32_t)__popcll(m & below);
      qEntry[wv][pos] = entry;
      qTag[wv][pos] = grp | (ci << 2);
    }
    qn += (uint32_t)__popcll(m);
    if (qn >= 64u) drain(64u);
  }
  if (qn > 0u) drain(qn);
  if (prod < nProd && sub == 0) S.datCount[prod] = count;
}

// The frame's demand counters keep the larger of its two queries (step
// start and accepted predictor pose); the working count of the stream in
// use is datStart[numParticles + numEdges].
__global__ void k_dat_finalize(DeviceState S, StepParams P) {
  uint32_t nProd = P.numParticles + P.numEdges;
  uint32_t total = S.datStart[nProd];
  uint32_t vt = S.datStart[P.numParticles];
  atomicMax(&S.counters[CTR_DAT_PAIRS], total);
  atomicMax(&S.counters[CTR_DAT_VT], vt);
  atomicMax(&S.counters[CTR_DAT_EE], total - vt);
}

// ---------------------------------------------------------------------------
// Planar-DAT truncation (27_planar_dat.metal:536-560, 291-337, 756-906,
// 1176-1261, 1289-1310).  Each pair of the stream places a division plane
// between its sides at the anchor pose; a vertex approaching it keeps the
// fraction of its displacement that stops short of it (relaxation 0.85,
// time margin 1e-3).  The reduction takes the minimum over a vertex's pairs
// with an unsigned atomic minimum on the f32 bits (fractions are in
// [0, 1], where bit order is value order), so the result does not depend on
// the schedule; the apply pass consumes the fraction, restores 1.0 and caps
// the displacement at 0.5 * 0.85 * rq.
// ---------------------------------------------------------------------------
namespace {

constexpr float kDatRelax = 0.85f, kDatTimeMargin = 1.0e-3f;
constexpr uint32_t kDatFullPass = 0xFFFFFFFFu;

__device__ __forceinline__ void dat_lower(const DeviceState& S, uint32_t v, float t) {
  if (!(t > 0.0f)) t = 0.0f;  // +0 for -0 (unsigned order)
  atomicMin(&S.datT[v], as_uint(t));
}

__device__ __forceinline__ float dat_plane_fraction(const DeviceState& S, V3 x, V3 dx, V3 n, V3 plane, float side,
                                                    bool report) {
  float s0 = side * dot(n, x - plane);
  float q = side * dot(n, dx);
  if (!finite_bits(s0) || !finite_bits(q)) {
    if (report) {
      atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
      atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
    }
    return 0.0f;
  }
  if (s0 < -kDatGeometryEps) return 0.0f;
  // an observed sign change is constrained however small the motion
  if (q >= 0.0f || s0 + q > 0.0f) return 1.0f;
  float hit = fmaxf(s0, 0.0f) / (-q);
  if (hit > 1.0f + kDatGeometryEps) return 1.0f;
  return clampf(fminf(kDatRelax * hit, hit - kDatTimeMargin), 0.0f, 1.0f);
}

__device__ __forceinline__ void dat_vertex(const DeviceState& S, uint32_t v, V3 x, V3 dx, V3 n, V3 plane,
                                           float side) {
  float t = dat_plane_fraction(S, x, dx, n, plane, side, true);
  if (t < 1.0f) dat_lower(S, v, t);
}

__device__ __forceinline__ bool dat_vt_fallback(V3 a, V3 b, V3 c, V3 rel, V3& n) {
  V3 raw = cross(b - a, c - a);
  float mag = sqrtf(dot(raw, raw));
  if (!(finite_bits(mag) && mag >= kDatGeometryEps)) return false;
  n = raw / mag;
  if (dot(n, rel) > 0.0f) n = -n;
  return true;
}

__device__ __forceinline__ bool dat_ee_fallback(V3 a0, V3 a1, V3 b0, V3 b1, V3 rel, V3& n) {
  V3 ua = a1 - a0, ub = b1 - b0;
  V3 raw = cross(ua, ub);
  float mag = sqrtf(dot(raw, raw));
  if (finite_bits(mag) && mag >= kDatGeometryEps) {
    n = raw / mag;
  } else {
    float la = sqrtf(dot(ua, ua)), lb = sqrtf(dot(ub, ub));
    V3 u = la >= lb ? ua : ub;
    float ul = fmaxf(la, lb);
    if (!finite_bits(ul) || ul < kDatGeometryEps) return false;
    u = u / ul;
    V3 perp = rel - u * dot(u, rel);
    float pl = sqrtf(dot(perp, perp));
    if (!finite_bits(pl) || pl < kDatGeometryEps) {
      V3 axis = fabsf(u.x) < 0.7f ? v3(1, 0, 0) : v3(0, 1, 0);
      perp = cross(u, axis);
      pl = sqrtf(dot(perp, perp));
    }
    if (!finite_bits(pl) || pl < kDatGeometryEps) return false;
    n = perp / pl;
  }
  if (dot(n, rel) > 0.0f) n = -n;
  return finite3(n);
}

__device__ __forceinline__ bool dat_active(const DeviceState& S, uint32_t b, uint32_t color) {
  return color == kDatFullPass || (S.posLin[b].w > 0.0f && S.colors[b] == color);
}

__device__ __forceinline__ void dat_freeze4(const DeviceState& S, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  dat_lower(S, a, 0.0f);
  dat_lower(S, b, 0.0f);
  dat_lower(S, c, 0.0f);
  dat_lower(S, d, 0.0f);
}

__device__ void dat_reduce_pair(const DeviceState& S, const StepParams& P, const F4* anchor, U4 pr, uint32_t color) {
  const float rq = P.datQueryRadius;
  if (pr.x == kDatVT) {
    const uint32_t v = pr.y;
    const U4 t = S.tris[pr.z];
    const bool aV = dat_active(S, v, color), aA = dat_active(S, t.x, color), aB = dat_active(S, t.y, color),
               aC = dat_active(S, t.z, color);
    if (!aV && !aA && !aB && !aC) return;
    V3 xv = xyz(anchor[v]), a = xyz(anchor[t.x]), b = xyz(anchor[t.y]), c = xyz(anchor[t.z]);
    V3 bary;
    V3 cp = closest_pt_triangle(xv, a, b, c, bary);
    V3 nh = xv - cp;
    float gap = sqrtf(dot(nh, nh));
    if (!finite_bits(gap)) {
      atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
      atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
      dat_freeze4(S, v, t.x, t.y, t.z);
      return;
    }
    if (gap > rq) return;
    V3 dv = P3(S, v) - xv, da = P3(S, t.x) - a, db = P3(S, t.y) - b, dc = P3(S, t.z) - c;
    V3 n;
    if (gap >= kDatGeometryEps) {
      n = nh / gap;
    } else {
      if (!dat_vt_fallback(a, b, c, dv - (da + db + dc) / 3.0f, n)) {
        atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
        atomicAdd(&S.counters[CTR_DAT_VT_DEGEN], 1u);
        dat_freeze4(S, v, t.x, t.y, t.z);
        return;
      }
      atomicAdd(&S.counters[CTR_DAT_VT_DEGEN], 1u);
    }
    float apV = fmaxf(-dot(n, dv), 0.0f);
    float apT = fmaxf(fmaxf(dot(n, da), dot(n, db)), fmaxf(dot(n, dc), 0.0f));
    float lambda = apV + apT == 0.0f ? 0.5f : clampf(apT / (apV + apT), 0.05f, 0.95f);
    V3 plane = gap >= kDatGeometryEps ? cp + nh * lambda : cp;
    if (apV > 0.0f && aV) dat_vertex(S, v, xv, dv, n, plane, 1.0f);
    if (apT > 0.0f) {
      if (aA) dat_vertex(S, t.x, a, da, n, plane, -1.0f);
      if (aB) dat_vertex(S, t.y, b, db, n, plane, -1.0f);
      if (aC) dat_vertex(S, t.z, c, dc, n, plane, -1.0f);
    }
    return;
  }
  const U2 ea = S.edges[pr.y], eb = S.edges[pr.z];
  const bool aA0 = dat_active(S, ea.x, color), aA1 = dat_active(S, ea.y, color), aB0 = dat_active(S, eb.x, color),
             aB1 = dat_active(S, eb.y, color);
  if (!aA0 && !aA1 && !aB0 && !aB1) return;
  V3 a0 = xyz(anchor[ea.x]), a1 = xyz(anchor[ea.y]), b0 = xyz(anchor[eb.x]), b1 = xyz(anchor[eb.y]);
  float s, t;
  closest_seg_seg(a0, a1, b0, b1, s, t);
  V3 ca = a0 + (a1 - a0) * s, cb = b0 + (b1 - b0) * t;
  V3 nh = ca - cb;
  float gap = sqrtf(dot(nh, nh));
  if (!finite_bits(gap)) {
    atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
    atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
    dat_freeze4(S, ea.x, ea.y, eb.x, eb.y);
    return;
  }
  if (gap > rq) return;
  V3 da0 = P3(S, ea.x) - a0, da1 = P3(S, ea.y) - a1, db0 = P3(S, eb.x) - b0, db1 = P3(S, eb.y) - b1;
  V3 n;
  if (gap > kDatGeometryEps) {
    n = nh / gap;
  } else {
    if (!dat_ee_fallback(a0, a1, b0, b1, (da0 + da1 - db0 - db1) * 0.5f, n)) {
      atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
      atomicAdd(&S.counters[CTR_DAT_EE_DEGEN], 1u);
      dat_freeze4(S, ea.x, ea.y, eb.x, eb.y);
      return;
    }
    atomicAdd(&S.counters[CTR_DAT_EE_DEGEN], 1u);
  }
  float apA = fmaxf(fmaxf(-dot(n, da0), -dot(n, da1)), 0.0f);
  float apB = fmaxf(fmaxf(dot(n, db0), dot(n, db1)), 0.0f);
  float lambda = apA + apB == 0.0f ? 0.5f : clampf(apB / (apA + apB), 0.05f, 0.95f);
  V3 plane = gap > kDatGeometryEps ? cb + nh * lambda : cb;
  if (apA > 0.0f) {
    if (aA0) dat_vertex(S, ea.x, a0, da0, n, plane, 1.0f);
    if (aA1) dat_vertex(S, ea.y, a1, da1, n, plane, 1.0f);
  }
  if (apB > 0.0f) {
    if (aB0) dat_vertex(S, eb.x, b0, db0, n, plane, -1.0f);
    if (aB1) dat_vertex(S, eb.y, b1, db1, n, plane, -1.0f);
  }
}

}  // namespace

// one thread per pair of the stream in use; `color` = kDatFullPass after
// the predictor, else the static color whose primal just ran
__global__ void __launch_bounds__(256) k_dat_reduce(DeviceState S, StepParams P, const F4* anchor, uint32_t color) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;  // the frame's residual check has ended its iterations
  const uint32_t total = S.datStart[P.numParticles + P.numEdges];
  const uint32_t n = total < P.maxDatPairs ? total : P.maxDatPairs;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dat_reduce_pair(S, P, anchor, S.datPairs[i], color);
}

// full pass: every surface particle (also refreshing the start-of-iteration
// snapshot); color pass: the color's bodies from colorWork[start, start+count)
__global__ void __launch_bounds__(256) k_dat_apply(DeviceState S, StepParams P, const F4* anchor, uint32_t full,
                                                   uint32_t start, uint32_t count) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v;
  if (full) {
    if (i >= P.numParticles) return;
    v = S.particleIdx[i];
  } else {
    if (i >= count) return;
    v = S.colorWork[start + i].x;
    if (S.clothGroup[v] == 0u) return;
  }
  const F4 cur = S.posLin[v];
  const V3 x0 = xyz(anchor[v]);
  V3 d = xyz(cur) - x0;
  const bool unsafeCap = S.datStart[P.numParticles + P.numEdges] > P.maxDatPairs ||
                         S.counters[CTR_DAT_PAIRS] > P.maxDatPairs || S.counters[CTR_DAT_OVERFLOW] > 0u ||
                         __hip_atomic_load(&S.counters[CTR_DAT_INVALID], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0u;
  float t = as_float(S.datT[v]);
  S.datT[v] = as_uint(1.0f);
  if (unsafeCap) t = 0.0f;
  if (!finite3(d) || !finite_bits(t)) {
    atomicAdd(&S.counters[CTR_DAT_INVALID], 1u);
    atomicAdd(&S.counters[CTR_DAT_NONFINITE], 1u);
    d = v3s(0.0f);
    t = 0.0f;
  }
  bool truncated = t < 1.0f;
  d = d * clampf(t, 0.0f, 1.0f);
  const float maxD = 0.5f * kDatRelax * P.datQueryRadius;
  const float d2 = dot(d, d);
  if (d2 > maxD * maxD) {
    d = d * (maxD / sqrtf(d2));
    truncated = true;
  }
  if (truncated) {
    const F4 np = f4(x0 + d, cur.w);
    S.posLin[v] = np;
    S.posLo[v] = F4{0.0f, 0.0f, 0.0f, 0.0f};
    if (full) S.posLinSnap[v] = np;
    atomicAdd(&S.counters[CTR_DAT_TRUNC], 1u);
  }
}

__global__ void __launch_bounds__(256) k_dat_reanchor(DeviceState S, StepParams P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.numParticles) return;
  const uint32_t v = S.particleIdx[i];
  S.datAnchor[v] = f4(xyz(S.posLin[v]), 0.0f);
}

// a failed surface query restores every particle to its step-start pose
__global__ void __launch_bounds__(256) k_dat_restore(DeviceState S, StepParams P) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.numBodies || S.shape[b].w >= 0.0f) return;
  const bool failed = S.counters[CTR_SOFT_CANDIDATES] > P.maxSoft || S.counters[CTR_DAT_PAIRS] > P.maxDatPairs ||
                      S.counters[CTR_DAT_OVERFLOW] > 0u || S.counters[CTR_DAT_INVALID] > 0u;
  if (failed) {
    S.posLin[b] = f4(xyz(S.initLin[b]), S.posLin[b].w);
    S.posLo[b] = F4{0.0f, 0.0f, 0.0f, 0.0f};
  }
}

// ---------------------------------------------------------------------------
// soft persistence map (25_cloth.metal:79-137)
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ uint32_t soft_key_hash(uint32_t a, uint32_t b, uint32_t cap) {
  uint32_t h = (a * 0x9E3779B1u) ^ (b * 0x85EBCA77u);
  h ^= h >> 16;
  h *= 0xC2B2AE3Du;
  h ^= h >> 13;
  return h & (cap - 1u);
}

__device__ int soft_map_find(const DeviceState& S, const StepParams& P, uint32_t keyA, uint32_t keyB) {
  uint32_t h = soft_key_hash(keyA, keyB, P.softMapCapacity);
  for (uint32_t probe = 0; probe < 64u; probe++) {
    uint32_t ka = __hip_atomic_load(&S.softMapKeyA[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ka == 0u) return -1;
    if (ka == keyA + 1u && S.softMapKeyB[h] == keyB) return (int)S.softMapVal[h];
    h = (h + 1u) & (P.softMapCapacity - 1u);
  }
  return -1;
}

// previous side (+1 / -1) of a persisting contact, else the proposal
__device__ float soft_prev_side(const DeviceState& S, const StepParams& P, uint32_t keyA, uint32_t keyB,
                                float proposed) {
  int prev = soft_map_find(S, P, keyA, keyB);
  if (prev < 0) return proposed;
  return (as_uint(S.prevSoft[prev].anchorA.w) & kSoftSideNeg) ? -1.0f : 1.0f;
}

__device__ __forceinline__ float soft_lambda_cap(const StepParams& P, float minMass) {
  return fminf(P.lambdaMax, fmaxf(10.0f, 1.0e5f * minMass));
}

// Warm start from the persistence map (tangential dual transported into the
// new basis), mass-aware penalty floor, record fill (25_cloth.metal:273-333).
__device__ SoftRec soft_record(const DeviceState& S, const StepParams& P, U4 ids, F4 weights, V3 n, float friction,
                               V3 anchorA, float C0n, float restT, float lamCap, float minMass, uint32_t kind,
                               bool rigidA, bool roundA, bool sideNeg, uint32_t keyA, uint32_t keyB) {
  V3 lam = v3s(0), pen = v3s(0);
  int prev = soft_map_find(S, P, keyA, keyB);
  if (prev >= 0) {
    const SoftRec& ps = S.prevSoft[prev];
    lam = xyz(ps.lambda) * (P.alpha * P.gamma);
    pen = xyz(ps.penalty);
    V3 t1o, t2o, t1n, t2n;
    orthonormal(xyz(ps.normal), t1o, t2o);
    orthonormal(n, t1n, t2n);
    V3 lt = t1o * lam.y + t2o * lam.z;
    lam.y = dot(t1n, lt);
    lam.z = dot(t2n, lt);
  }
  float penFloor = fmaxf(kPenaltyMin, minMass / (P.dt * P.dt));
  pen = v3(clampf(pen.x * P.gamma, penFloor, kPenaltyMax), clampf(pen.y * P.gamma, penFloor, kPenaltyMax),
           clampf(pen.z * P.gamma, penFloor, kPenaltyMax));
  uint32_t flags = (rigidA ? kSoftRigidA : 0u) | (sideNeg ? kSoftSideNeg : 0u) | (kind << kSoftKindShift) |
                   (roundA ? kSoftRoundA : 0u);
  SoftRec r;
  r.ids = ids;
  r.normal = f4(n, friction);
  r.anchorA = f4(rigidA ? anchorA : v3(restT, 0, 0), as_float(flags));
  r.weights = weights;
  r.C0 = F4{C0n, 0.0f, 0.0f, lamCap};
  r.lambda = f4(lam, as_float(keyA));
  r.penalty = f4(pen, as_float(keyB));
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// rigid feature points vs triangles at the step-start pose
// (25_cloth.metal:928-1142): box corners / sphere centres of the globals and
// hashed rigid colliders against every triangle, at most 4 per triangle
// ---------------------------------------------------------------------------
namespace {

struct RtFeature {
  V3 world;
  float radius;
  uint32_t id;
};

__device__ int rt_features(const DeviceState& S, uint32_t c, V3 center, F4 rot, V3 m, float rT, float margin,
                           RtFeature* out) {
  uint32_t kind = S.colliderShapeType[c] & kShapeKindMask;
  V3 size = xyz(S.colliderShape[c]);
  if (kind == 1u) {
    out[0] = RtFeature{center, size.x * 0.5f, 8u};
    return 1;
  }
  if (kind == 3u) {
    // capsule: spheres at its two ends and middle (25_cloth.metal:957-970)
    const float half = size.x * 0.5f, rc = size.y;
    const V3 ax = q_rotate(rot, v3(0.0f, 0.0f, 1.0f));
    int n = 0;
    for (int i = -1; i <= 1; i++) {
      const V3 w = center + ax * (half * (float)i);
      if (length(w - m) > rT + rc + 4.0f * margin) continue;
      out[n++] = RtFeature{w, rc, (uint32_t)(10 + i)};
    }
    return n;
  }
  if (kind != 0u) return 0;  // torus: its particle contacts come from the rigid narrow phase
  V3 h3 = size * 0.5f;
  int n = 0;
  for (uint32_t i = 0; i < 8u; i++) {
    V3 local = v3((i & 1u) ? h3.x : -h3.x, (i & 2u) ? h3.y : -h3.y, (i & 4u) ? h3.z : -h3.z);
    V3 w = xform(center, rot, local);
    if (length(w - m) > rT + 4.0f * margin) continue;
    if (n < 4) out[n++] = RtFeature{w, 0.0f, i};
  }
  return n;
}

// One (collider, triangle) test; emits through f(record) up to the cap.
template <typename F>
__device__ void rt_test(const DeviceState& S, const StepParams& P, uint32_t tri, U4 tid, V3 a, V3 b, V3 c, V3 m,
                        float rT, float rt_, V3 triN, V3 velT, uint32_t col, uint32_t& emitted, F&& f) {
  uint32_t o = S.colliderOwner[col];
  F4 bq = S.posAng[o];
  V3 bp = P3(S, o);
  V3 center = bp + q_rotate(bq, xyz(S.colliderLocalPos[col]));
  F4 rot = q_mul(bq, S.colliderLocalRot[col]);
  float rb = fabsf(S.colliderShape[col].w);
  bool nearO = length(center - m) <= rT + rb + P.collisionMargin;
  uint32_t kind = S.colliderShapeType[col] & kShapeKindMask;
  if (nearO && kind == 0u) {
    V3 lm = q_rotate(q_conj(rot), m - center);
    V3 dq2 = vmax(vabs(lm) - xyz(S.colliderShape[col]) * 0.5f, v3s(0.0f));
    nearO = length(dq2) <= rT + P.collisionMargin + 0.05f;
  }
  if (!nearO) return;
  if (pair_excluded(S, P.numExclusions, min(o, tid.x), max(o, tid.x)) ||
      pair_excluded(S, P.numExclusions, min(o, tid.y), max(o, tid.y)) ||
      pair_excluded(S, P.numExclusions, min(o, tid.z), max(o, tid.z)))
    return;
  RtFeature feats[4];
  int nf = rt_features(S, col, center, rot, m, rT, P.collisionMargin, feats);
  for (int fi = 0; fi < nf && emitted < kRtMaxPerTri; fi++) {
    V3 bary;
    V3 q = closest_pt_triangle(feats[fi].world, a, b, c, bary);
    V3 d = feats[fi].world - q;
    float dist = length(d);
    float hSum = rt_ + feats[fi].radius;
    float inflate = fminf(length(xyz(S.velLin[o]) - velT) * P.dt, 0.5f * rT);
    if (dist - hSum > P.elemMargin + inflate) continue;
    float side = dot(d, triN) >= 0.0f ? 1.0f : -1.0f;
    uint32_t keyA = (kSoftRT << 28) | col;
    uint32_t keyB = tri * 16u + feats[fi].id;
    float proposed = dot(bp - q, triN) >= 0.0f ? 1.0f : -1.0f;
    bool interior = bary.x > 0.02f && bary.y > 0.02f && bary.z > 0.02f;
    float sMem = interior ? soft_prev_side(S, P, keyA, keyB, proposed) : proposed;
    V3 n;
    float g;
    if (dist > 1e-7f && side == sMem) { n = d / dist; g = dist; }
    else { n = triN * sMem; g = -dist; }
    float fricT = (S.props[tid.x].w + S.props[tid.y].w + S.props[tid.z].w) / 3.0f;
    float friction = sqrtf(fmaxf(S.props[o].w * fricT, 0.0f));
    float mo = S.posLin[o].w;
    float minMass = mo > 0.0f ? mo : kFltMax;
    float mA = S.posLin[tid.x].w, mB = S.posLin[tid.y].w, mC = S.posLin[tid.z].w;
    if (mA > 0.0f) minMass = fminf(minMass, mA);
    if (mB > 0.0f) minMass = fminf(minMass, mB);
    if (mC > 0.0f) minMass = fminf(minMass, mC);
    if (minMass == kFltMax) continue;
    bool roundA = kind != 0u;
    V3 anchor = roundA ? feats[fi].world - bp : q_rotate(q_conj(bq), feats[fi].world - bp);
    f(o, tid, bary, n, friction, anchor, g - hSum + P.elemMargin, hSum - P.elemMargin, minMass, roundA,
      sMem < 0.0f, keyA, keyB);
    emitted++;
  }
}

}  // namespace

template <bool EMIT>
__global__ void __launch_bounds__(256) k_rt_emit(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (!EMIT && gid == 0) S.softFlag[P.numTris] = 0u;
  if (gid >= P.numTris) return;
  U4 tid = S.tris[gid];
  V3 a = P3(S, tid.x), b = P3(S, tid.y), c = P3(S, tid.z);
  V3 m = (a + b + c) / 3.0f;
  float rt_ = fmaxf(fabsf(S.shape[tid.x].w), fmaxf(fabsf(S.shape[tid.y].w), fabsf(S.shape[tid.z].w)));
  float rT = fmaxf(length(m - a), fmaxf(length(m - b), length(m - c))) + rt_;
  V3 triN = cross(b - a, c - a);
  float tnLen = length(triN);
  uint32_t emitted = 0;
  uint32_t base = EMIT ? S.softStart[gid] : 0u;
  if (tnLen >= 1e-12f) {
    triN = triN / tnLen;
    V3 velT = (xyz(S.velLin[tid.x]) + xyz(S.velLin[tid.y]) + xyz(S.velLin[tid.z])) / 3.0f;
    auto emit = [&](uint32_t o, U4 t, V3 bary, V3 n, float friction, V3 anchor, float C0n, float restT,
                    float minMass, bool roundA, bool sideNeg, uint32_t keyA, uint32_t keyB) {
      if (!EMIT) return;
      uint32_t slot = base + emitted;
      if (slot >= P.maxSoft) return;
      S.soft[slot] = soft_record(S, P, U4{o, t.x, t.y, t.z}, F4{1.0f, -bary.x, -bary.y, -bary.z}, n, friction,
                                 anchor, C0n, restT, soft_lambda_cap(P, minMass), minMass, kSoftRT, true, roundA,
                                 sideNeg, keyA, keyB);
    };
    if (P.numHashedRigid > 0u && P.numHashedRigid <= 8u) {
      for (uint32_t ri = 0; ri < P.numHashedRigid && emitted < kRtMaxPerTri; ri++)
        rt_test(S, P, gid, tid, a, b, c, m, rT, rt_, triN, velT, S.hashedRigid[ri], emitted, emit);
    } else if (P.numHashedRigid > 0u) {
      // widen the scan so centres further than one cell are still reached
      int R = (int)ceilf((rT + 0.5f * P.cellSize) / P.cellSize);
      R = R < 1 ? 1 : (R > 4 ? 4 : R);
      int cx, cy, cz;
      cell_coord(m, P.cellSize, cx, cy, cz);
      for (int dz = -R; dz <= R && emitted < kRtMaxPerTri; dz++)
        for (int dy = -R; dy <= R && emitted < kRtMaxPerTri; dy++)
          for (int dx = -R; dx <= R && emitted < kRtMaxPerTri; dx++) {
            uint32_t h = cell_hash(cx + dx, cy + dy, cz + dz, 0u, P.gridHashSize);
            uint32_t s = S.cellStart[h], e = s + S.cellCount[h];
            for (uint32_t k = s; k < e && emitted < kRtMaxPerTri; k++) {
              uint32_t col = S.cellBodies[k];
              if (S.colliderShape[col].w < 0.0f) continue;  // particles: V-T covers them
              rt_test(S, P, gid, tid, a, b, c, m, rT, rt_, triN, velT, col, emitted, emit);
            }
          }
    }
    for (uint32_t gI = 0; gI < P.numGlobals && emitted < kRtMaxPerTri; gI++) {
      uint32_t col = S.globalIdx[gI];
      if (S.colliderShape[col].w < 0.0f) continue;
      rt_test(S, P, gid, tid, a, b, c, m, rT, rt_, triN, velT, col, emitted, emit);
    }
  }
  if (!EMIT) S.softFlag[gid] = emitted;
  else if (gid == P.numTris - 1u) S.counters[CTR_SOFT_RT] = S.softStart[P.numTris];
}

// ---------------------------------------------------------------------------
// OGC contacts from the accepted pair stream at the predictor pose, C0
// back-evaluated to the step-start reference (27_planar_dat.metal:344-510)
// ---------------------------------------------------------------------------
namespace {

__device__ bool dat_contact(const DeviceState& S, const StepParams& P, U4 pair, SoftRec& out) {
  const F4* ref = S.initLin;
  if (pair.x == kDatVT) {
    uint32_t v = pair.y, ti = pair.z;
    if (ti >= P.numTris) return false;
    U4 t = S.tris[ti];
    V3 p = P3(S, v), a = P3(S, t.x), b = P3(S, t.y), c = P3(S, t.z);
    V3 bary;
    V3 q = closest_pt_triangle(p, a, b, c, bary);
    V3 d = p - q;
    float dist2 = dot(d, d);
    float rv = fabsf(S.shape[v].w);
    float rt = fmaxf(fabsf(S.shape[t.x].w), fmaxf(fabsf(S.shape[t.y].w), fabsf(S.shape[t.z].w)));
    float hSum = rv + rt;
    V3 velT = (xyz(S.velLin[t.x]) + xyz(S.velLin[t.y]) + xyz(S.velLin[t.z])) / 3.0f;
    float inflate = fminf(length(xyz(S.velLin[v]) - velT) * P.dt, 0.3f * P.surfaceContactCellSize);
    float detect = fminf(hSum + P.elemMargin + inflate, P.datQueryRadius);
    if (dist2 > detect * detect) return false;
    float dist = sqrtf(dist2);
    V3 triN = cross(b - a, c - a);
    float nl = length(triN);
    if (nl < kDatGeometryEps) return false;
    triN = triN / nl;
    float side = dot(d, triN) >= 0.0f ? 1.0f : -1.0f;
    uint32_t keyA = (kSoftVT << 28) | v, keyB = ti;
    bool interior = bary.x > 0.02f && bary.y > 0.02f && bary.z > 0.02f;
    float remembered = interior ? soft_prev_side(S, P, keyA, keyB, side) : side;
    V3 n;
    float gap;
    if (dist > 1.0e-7f && side == remembered) { n = d / dist; gap = dist; }
    else { n = triN * remembered; gap = -dist; }
    float fricT = S.props[t.x].w * bary.x + S.props[t.y].w * bary.y + S.props[t.z].w * bary.z;
    float friction = sqrtf(fmaxf(S.props[v].w * fricT, 0.0f));
    float mv = S.posLin[v].w;
    float minMass = mv > 0.0f ? mv : kFltMax;
    uint32_t ids[3] = {t.x, t.y, t.z};
    for (int j = 0; j < 3; j++) {
      float mj = S.posLin[ids[j]].w;
      if (mj > 0.0f) minMass = fminf(minMass, mj);
    }
    if (minMass == kFltMax) return false;
    V3 dq = (p - xyz(ref[v])) - (a - xyz(ref[t.x])) * bary.x - (b - xyz(ref[t.y])) * bary.y -
            (c - xyz(ref[t.z])) * bary.z;
    float Cdetect = gap - hSum + P.elemMargin;
    float C0n = Cdetect - dot(n, dq);
    out = soft_record(S, P, U4{v, t.x, t.y, t.z}, F4{1.0f, -bary.x, -bary.y, -bary.z}, n, friction, v3s(0), C0n,
                      hSum - P.elemMargin, soft_lambda_cap(P, minMass), minMass, kSoftVT, false, false,
                      remembered < 0.0f, keyA, keyB);
    return true;
  }
  uint32_t ea = pair.y, eb = pair.z;
  if (ea >= P.numEdges || eb >= P.numEdges) return false;
  U2 ai = S.edges[ea], bi = S.edges[eb];
  // E-E contacts between authored edges, or within a self-colliding
  // surface; other pairs (tet-boundary edges) only bound the motion
  // (27_planar_dat.metal:196-201)
  if (ea >= P.numContactEdges || eb >= P.numContactEdges) {
    const uint32_t g = S.clothGroup[ai.x];
    const bool same = g != 0u && S.clothGroup[ai.y] == g && S.clothGroup[bi.x] == g && S.clothGroup[bi.y] == g;
    if (!(same && S.selfCollision[ai.x] != 0u)) return false;
  }
  V3 a0 = P3(S, ai.x), a1 = P3(S, ai.y), b0 = P3(S, bi.x), b1 = P3(S, bi.y);
  float s, t;
  closest_seg_seg(a0, a1, b0, b1, s, t);
  if (s < kEeInteriorEps || s > 1.0f - kEeInteriorEps || t < kEeInteriorEps || t > 1.0f - kEeInteriorEps)
    return false;
  V3 ca = a0 + (a1 - a0) * s, cb = b0 + (b1 - b0) * t;
  V3 d = ca - cb;
  float dist = length(d);
  float rA = fmaxf(fabsf(S.shape[ai.x].w), fabsf(S.shape[ai.y].w));
  float rB = fmaxf(fabsf(S.shape[bi.x].w), fabsf(S.shape[bi.y].w));
  float hSum = rA + rB;
  V3 va = (xyz(S.velLin[ai.x]) + xyz(S.velLin[ai.y])) * 0.5f;
  V3 vb = (xyz(S.velLin[bi.x]) + xyz(S.velLin[bi.y])) * 0.5f;
  float inflate = fminf(length(va - vb) * P.dt, 0.3f * P.surfaceContactCellSize);
  float detect = fminf(hSum + P.elemMargin + inflate, P.datQueryRadius);
  if (dist > detect) return false;
  uint32_t keyA = (kSoftEE << 28) | ea, keyB = eb;
  int previous = soft_map_find(S, P, keyA, keyB);
  V3 n;
  float gap = dist;
  if (finite_bits(dist) && dist >= 1.0e-7f) {
    V3 nGeo = d / dist;
    n = nGeo;
    if (previous >= 0 && dot(nGeo, xyz(S.prevSoft[previous].normal)) < 0.0f) {
      n = -nGeo;
      gap = -dist;
    }
  } else if (previous >= 0 && finite3(xyz(S.prevSoft[previous].normal))) {
    n = xyz(S.prevSoft[previous].normal);
    gap = 0.0f;
  } else {
    // fallback normal from the edge directions / relative motion
    V3 ua = a1 - a0, ub = b1 - b0, rel = va - vb;
    V3 raw = cross(ua, ub);
    float mag = length(raw);
    if (finite_bits(mag) && mag >= kDatGeometryEps) {
      n = raw / mag;
    } else {
      float la = length(ua), lb = length(ub);
      V3 u = la >= lb ? ua : ub;
      float ul = fmaxf(la, lb);
      if (!finite_bits(ul) || ul < kDatGeometryEps) return false;
      u = u / ul;
      V3 perp = rel - u * dot(u, rel);
      float pl = length(perp);
      if (!finite_bits(pl) || pl < kDatGeometryEps) {
        V3 axis = fabsf(u.x) < 0.7f ? v3(1, 0, 0) : v3(0, 1, 0);
        perp = cross(u, axis);
        pl = length(perp);
      }
      if (!finite_bits(pl) || pl < kDatGeometryEps) return false;
      n = perp / pl;
    }
    if (dot(n, rel) > 0.0f) n = -n;
    if (!finite3(n)) return false;
    gap = 0.0f;
  }
  float friction = sqrtf(fmaxf((S.props[ai.x].w + S.props[ai.y].w) * 0.5f * (S.props[bi.x].w + S.props[bi.y].w) * 0.5f,
                               0.0f));
  float minMass = kFltMax;
  uint32_t ids[4] = {ai.x, ai.y, bi.x, bi.y};
  for (int j = 0; j < 4; j++) {
    float mj = S.posLin[ids[j]].w;
    if (mj > 0.0f) minMass = fminf(minMass, mj);
  }
  if (minMass == kFltMax) return false;
  V3 dq = (a0 - xyz(ref[ai.x])) * (1.0f - s) + (a1 - xyz(ref[ai.y])) * s - (b0 - xyz(ref[bi.x])) * (1.0f - t) -
          (b1 - xyz(ref[bi.y])) * t;
  float Cdetect = gap - hSum + P.elemMargin;
  float C0n = Cdetect - dot(n, dq);
  out = soft_record(S, P, U4{ai.x, ai.y, bi.x, bi.y}, F4{1.0f - s, s, -(1.0f - t), -t}, n, friction, v3s(0), C0n,
                    hSum - P.elemMargin, soft_lambda_cap(P, minMass), minMass, kSoftEE, false, false, false, keyA,
                    keyB);
  return true;
}

}  // namespace

template <bool EMIT>
__global__ void __launch_bounds__(256) k_dat_contacts(DeviceState S, StepParams P) {
  uint32_t total = S.datStart[P.numParticles + P.numEdges];
  uint32_t n = total < P.maxDatPairs ? total : P.maxDatPairs;
  uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (!EMIT && i0 == 0) S.softFlag[n] = 0u;
  uint32_t rt = EMIT ? S.counters[CTR_SOFT_RT] : 0u;
  for (uint32_t i = i0; i < n; i += gridDim.x * blockDim.x) {
    SoftRec r;
    bool ok = dat_contact(S, P, S.datPairs[i], r);
    if (!EMIT) {
      S.softFlag[i] = ok ? 1u : 0u;
    } else if (ok) {
      uint32_t slot = rt + S.softStart[i];
      if (slot < P.maxSoft) S.soft[slot] = r;
    }
  }
}

// total soft demand; clamp the working count (25_cloth.metal:1149-1163)
__global__ void k_soft_finalize(DeviceState S, StepParams P) {
  uint32_t total = P.pairStream ? S.datStart[P.numParticles + P.numEdges] : 0u;
  uint32_t n = total < P.maxDatPairs ? total : P.maxDatPairs;
  uint32_t demand = S.counters[CTR_SOFT_RT] + (P.pairStream ? S.softStart[n] : 0u) +
                    (P.isoOn ? S.counters[CTR_SOFT_ISO] : 0u);
  S.counters[CTR_SOFT_CANDIDATES] = demand;
  S.counters[CTR_SOFT] = demand < P.maxSoft ? demand : P.maxSoft;
}

// debug lookup: (keyA, keyB) -> index of the previous frame's record, or -1
__global__ void __launch_bounds__(256) k_softmap_query(DeviceState S, StepParams P, const U2* q, int32_t* out,
                                                       uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = soft_map_find(S, P, q[i].x, q[i].y);
}

__global__ void __launch_bounds__(256) k_softmap_clear(DeviceState S, StepParams P) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.softMapCapacity; i += gridDim.x * blockDim.x)
    S.softMapKeyA[i] = 0u;
}

__global__ void __launch_bounds__(256) k_softmap_insert(DeviceState S, StepParams P) {
  uint32_t n = S.counters[CTR_SOFT];
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    uint32_t keyA = as_uint(S.soft[g].lambda.w), keyB = as_uint(S.soft[g].penalty.w);
    uint32_t h = soft_key_hash(keyA, keyB, P.softMapCapacity);
    for (uint32_t probe = 0; probe < P.softMapCapacity; probe++) {
      if (atomicCAS(&S.softMapKeyA[h], 0u, keyA + 1u) == 0u) {
        S.softMapKeyB[h] = keyB;
        S.softMapVal[h] = g;
        break;
      }
      h = (h + 1u) & (P.softMapCapacity - 1u);
    }
  }
}

// ---------------------------------------------------------------------------
// isotropic OGC limiter (25_cloth.metal:350-940): every surface particle
// tracks its four closest collision triangles and every authored edge its
// four closest edges across frames.  Per frame a small candidate pool is
// evaluated -- the tracked set, the triangles adjacent to the best, the mesh
// neighbours' best -- and the element grid is scanned on staggered reseeds
// or fast relative motion.  Each observed distance between elements outside
// each other's 2-ring lowers the squared clearance bound of the vertices
// involved (atomic min on the f32 bits: non-negative floats order as
// integers); the contacts come from the tracked sets.  Every tracker reads
// the previous frame's tracked sets (vtPrev / eePrev, not the ones being
// rewritten) and ties in distance keep the lower element id, so the result
// is a function of the inputs only.  Producers are the surface particles,
// then the authored edges; the records follow the rigid-feature ones in
// producer order.
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kOgcFarBits = 0x7F7F7F7Fu;  // cleared bound, 3.3961514e38

struct Best4 {
  uint32_t id0, id1, id2, id3;
  float d0, d1, d2, d3;
};
__device__ __forceinline__ void b4_init(Best4& b) {
  b.id0 = b.id1 = b.id2 = b.id3 = kNone;
  b.d0 = b.d1 = b.d2 = b.d3 = kFltMax;
}
__device__ __forceinline__ bool b4_has(const Best4& b, uint32_t t) {
  return b.id0 == t || b.id1 == t || b.id2 == t || b.id3 == t;
}
// (d, id) strictly before (e, jd); a NaN distance is before nothing
__device__ __forceinline__ bool b4_less(float d, uint32_t id, float e, uint32_t jd) {
  return d < e || (d == e && id < jd);
}
// keeps the four smallest (d2, id): the candidate replaces the last entry
// and sinks by compare-and-swap (fixed slots, so the set stays in registers)
__device__ __forceinline__ void b4_insert(Best4& b, uint32_t t, float d2) {
  if (!b4_less(d2, t, b.d3, b.id3)) return;
  b.id3 = t;
  b.d3 = d2;
  if (!b4_less(b.d3, b.id3, b.d2, b.id2)) return;
  { uint32_t i = b.id2; float d = b.d2; b.id2 = b.id3; b.d2 = b.d3; b.id3 = i; b.d3 = d; }
  if (!b4_less(b.d2, b.id2, b.d1, b.id1)) return;
  { uint32_t i = b.id1; float d = b.d1; b.id1 = b.id2; b.d1 = b.d2; b.id2 = i; b.d2 = d; }
  if (!b4_less(b.d1, b.id1, b.d0, b.id0)) return;
  { uint32_t i = b.id0; float d = b.d0; b.id0 = b.id1; b.d0 = b.d1; b.id1 = i; b.d1 = d; }
}
__device__ __forceinline__ uint32_t u4_at(U4 u, int k) { return k == 0 ? u.x : (k == 1 ? u.y : (k == 2 ? u.z : u.w)); }

__device__ __forceinline__ float contact_pad(const DeviceState& S, const StepParams& P, uint32_t a, uint32_t b) {
  return fmaxf(fabsf(S.shape[a].w), fabsf(S.shape[b].w)) + P.elemMargin;
}

// The 27-cell reseed and the bounds cap use the contact cell of the mode.
__device__ __forceinline__ bool iso_reseed(const StepParams& P, uint32_t id, uint32_t frame, V3 relV, bool missing,
                                           bool solid) {
  const float vth = 0.25f * P.surfaceContactCellSize / fmaxf(P.dt, 1e-6f);
  const uint32_t period = dot(relV, relV) > vth * vth ? 1u : 7u;
  return ((id + frame) & period) == 0u || (missing && !solid) || (missing && solid && ((id + frame) & 3u) == 0u);
}

struct VtQuery {
  uint32_t v, gv, ns, ne, n2s, n2e;
  bool solid;
  V3 p;
  float cap2;
};

// Solid key of an element: the closed surface's group when all its vertices
// lie on that surface and none on an authored triangle, else all ones.  A
// solid query skips exactly the candidates whose key is its own group, so a
// bucket whose least and greatest key are both that group is skipped whole.
__device__ __forceinline__ uint32_t solid_key3(const DeviceState& S, uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t g = S.clothGroup[a];
  return (g != 0u && S.clothGroup[b] == g && S.clothGroup[c] == g && S.clothVert[a] == 0u && S.clothVert[b] == 0u &&
          S.clothVert[c] == 0u)
             ? g
             : kNone;
}
__device__ __forceinline__ uint32_t solid_key2(const DeviceState& S, uint32_t a, uint32_t b) {
  const uint32_t g = S.clothGroup[a];
  return (g != 0u && S.clothGroup[b] == g && S.clothVert[a] == 0u && S.clothVert[b] == 0u) ? g : kNone;
}
__device__ __forceinline__ bool own_solid_bucket(const DeviceState& S, uint32_t slot, bool solid, uint32_t g) {
  return solid && S.isoSlotLo[slot] == g && S.isoSlotHi[slot] == g;
}

// one candidate triangle of a V-T tracker (25_cloth.metal:423-463)
__device__ __forceinline__ void vt_consider(const DeviceState& S, const StepParams& P, const VtQuery& q, Best4& best,
                                            float& own, uint32_t t) {
  if (t >= P.numTris || b4_has(best, t)) return;
  const U4 tv = S.tris[t];
  if (tv.x == q.v || tv.y == q.v || tv.z == q.v) return;
  if (q.solid && S.clothGroup[tv.x] == q.gv && S.clothGroup[tv.y] == q.gv && S.clothGroup[tv.z] == q.gv &&
      S.clothVert[tv.x] == 0u && S.clothVert[tv.y] == 0u && S.clothVert[tv.z] == 0u)
    return;
  V3 bary;
  const V3 c = closest_pt_triangle(q.p, P3(S, tv.x), P3(S, tv.y), P3(S, tv.z), bary);
  const V3 d = q.p - c;
  const float d2 = dot(d, d);
  if (!nbr_contains(S.nbrList, q.ns, q.ne, tv.x) && !nbr_contains(S.nbrList, q.ns, q.ne, tv.y) &&
      !nbr_contains(S.nbrList, q.ns, q.ne, tv.z))
    b4_insert(best, t, d2);
  if (d2 < q.cap2 && !nbr_contains(S.nbr2List, q.n2s, q.n2e, tv.x) &&
      !nbr_contains(S.nbr2List, q.n2s, q.n2e, tv.y) && !nbr_contains(S.nbr2List, q.n2s, q.n2e, tv.z)) {
    own = fminf(own, d2);
    const uint32_t bits = as_uint(d2);
    atomicMin(&S.ogcBound[tv.x], bits);
    atomicMin(&S.ogcBound[tv.y], bits);
    atomicMin(&S.ogcBound[tv.z], bits);
  }
}

struct EeQuery {
  uint32_t e, ga, ns0, ne0, ns1, ne1, n2s, n2e;
  U2 a;
  bool solid;
  V3 a0, a1;
  float cap2;
};

// one candidate edge of an E-E tracker (25_cloth.metal:697-737)
__device__ __forceinline__ void ee_consider(const DeviceState& S, const StepParams& P, const EeQuery& q, Best4& best,
                                            float& own, uint32_t o) {
  if (o >= P.numContactEdges || o == q.e || b4_has(best, o)) return;
  const U2 b = S.edges[o];
  if (b.x == q.a.x || b.x == q.a.y || b.y == q.a.x || b.y == q.a.y) return;
  if (q.solid && S.clothGroup[b.x] == q.ga && S.clothGroup[b.y] == q.ga && S.clothVert[b.x] == 0u &&
      S.clothVert[b.y] == 0u)
    return;
  float s, t;
  const V3 b0 = P3(S, b.x), b1 = P3(S, b.y);
  closest_seg_seg(q.a0, q.a1, b0, b1, s, t);
  const V3 d = (q.a0 + (q.a1 - q.a0) * s) - (b0 + (b1 - b0) * t);
  const float d2 = dot(d, d);
  if (!nbr_contains(S.nbrList, q.ns0, q.ne0, b.x) && !nbr_contains(S.nbrList, q.ns0, q.ne0, b.y) &&
      !nbr_contains(S.nbrList, q.ns1, q.ne1, b.x) && !nbr_contains(S.nbrList, q.ns1, q.ne1, b.y))
    b4_insert(best, o, d2);
  if (d2 < q.cap2 && !nbr_contains(S.nbr2List, q.n2s, q.n2e, b.x) && !nbr_contains(S.nbr2List, q.n2s, q.n2e, b.y)) {
    own = fminf(own, d2);
    const uint32_t bits = as_uint(d2);
    atomicMin(&S.ogcBound[b.x], bits);
    atomicMin(&S.ogcBound[b.y], bits);
  }
}

// V-T contacts of particle v's tracked triangles (25_cloth.metal:532-629);
// returns how many it emits, written from soft[base] when EMIT
template <bool EMIT>
__device__ uint32_t vt_emit(const DeviceState& S, const StepParams& P, uint32_t v, U4 tracked, uint32_t base) {
  const V3 pv = P3(S, v);
  const float rv = fabsf(S.shape[v].w), mv = S.posLin[v].w;
  const V3 velV = xyz(S.velLin[v]);
  uint32_t n = 0;
#pragma unroll
  for (int bi = 0; bi < 4; bi++) {
    const uint32_t t = u4_at(tracked, bi);
    if (t >= P.numTris) break;
    const U4 tv = S.tris[t];
    const V3 a = P3(S, tv.x), b = P3(S, tv.y), c = P3(S, tv.z);
    V3 bary;
    const V3 q = closest_pt_triangle(pv, a, b, c, bary);
    const V3 d = pv - q;
    const float dist2 = dot(d, d);
    const float rt = fmaxf(fabsf(S.shape[tv.x].w), fmaxf(fabsf(S.shape[tv.y].w), fabsf(S.shape[tv.z].w)));
    const float hSum = rv + rt;
    const V3 velT = (xyz(S.velLin[tv.x]) + xyz(S.velLin[tv.y]) + xyz(S.velLin[tv.z])) / 3.0f;
    const float inflate = fminf(length(velV - velT) * P.dt, 0.3f * P.surfaceContactCellSize);
    const float detect = hSum + P.elemMargin + inflate;
    if (dist2 > detect * detect) continue;
    V3 triN = cross(b - a, c - a);
    const float tnLen = length(triN);
    if (tnLen < 1e-12f) continue;
    float minMass = mv > 0.0f ? mv : kFltMax;
    const float mA = S.posLin[tv.x].w, mB = S.posLin[tv.y].w, mC = S.posLin[tv.z].w;
    if (mA > 0.0f) minMass = fminf(minMass, mA);
    if (mB > 0.0f) minMass = fminf(minMass, mB);
    if (mC > 0.0f) minMass = fminf(minMass, mC);
    if (minMass == kFltMax) continue;
    if (EMIT) {
      const uint32_t slot = base + n;
      if (slot < P.maxSoft) {
        const float dist = sqrtf(dist2);
        triN = triN / tnLen;
        const float side = dot(d, triN) >= 0.0f ? 1.0f : -1.0f;
        const uint32_t keyA = (kSoftVT << 28) | v, keyB = t;
        const bool deep = bary.x > 0.02f && bary.y > 0.02f && bary.z > 0.02f;
        const float mem = deep ? soft_prev_side(S, P, keyA, keyB, side) : side;
        V3 nrm;
        float g;
        if (dist > 1e-7f && side == mem) { nrm = d / dist; g = dist; }
        else { nrm = triN * mem; g = -dist; }
        const float fricT = S.props[tv.x].w * bary.x + S.props[tv.y].w * bary.y + S.props[tv.z].w * bary.z;
        const float friction = sqrtf(fmaxf(S.props[v].w * fricT, 0.0f));
        S.soft[slot] = soft_record(S, P, U4{v, tv.x, tv.y, tv.z}, F4{1.0f, -bary.x, -bary.y, -bary.z}, nrm, friction,
                                   v3s(0), g - hSum + P.elemMargin, hSum - P.elemMargin, soft_lambda_cap(P, minMass),
                                   minMass, kSoftVT, false, false, mem < 0.0f, keyA, keyB);
      }
    }
    n++;
  }
  return n;
}

// E-E contacts of edge e's tracked edges, each pair once by its lower edge
// (25_cloth.metal:789-853)
template <bool EMIT>
__device__ uint32_t ee_emit(const DeviceState& S, const StepParams& P, uint32_t e, U4 tracked, uint32_t base) {
  const U2 ea = S.edges[e];
  const V3 a0 = P3(S, ea.x), a1 = P3(S, ea.y);
  const float rA = fmaxf(fabsf(S.shape[ea.x].w), fabsf(S.shape[ea.y].w));
  const V3 velA = (xyz(S.velLin[ea.x]) + xyz(S.velLin[ea.y])) * 0.5f;
  uint32_t n = 0;
#pragma unroll
  for (int bi = 0; bi < 4; bi++) {
    const uint32_t o = u4_at(tracked, bi);
    if (o >= P.numContactEdges) break;
    if (o <= e) continue;
    const U2 eb = S.edges[o];
    const V3 b0 = P3(S, eb.x), b1 = P3(S, eb.y);
    float s, t;
    closest_seg_seg(a0, a1, b0, b1, s, t);
    if (s < kEeInteriorEps || s > 1.0f - kEeInteriorEps || t < kEeInteriorEps || t > 1.0f - kEeInteriorEps) continue;
    const V3 d = (a0 + (a1 - a0) * s) - (b0 + (b1 - b0) * t);
    const float dist = length(d);
    const float rB = fmaxf(fabsf(S.shape[eb.x].w), fabsf(S.shape[eb.y].w));
    const float hSum = rA + rB;
    const V3 velB = (xyz(S.velLin[eb.x]) + xyz(S.velLin[eb.y])) * 0.5f;
    const float inflate = fminf(length(velA - velB) * P.dt, 0.3f * P.surfaceContactCellSize);
    if (dist - hSum > P.elemMargin + inflate) continue;
    if (dist < 1e-7f) continue;
    float minMass = kFltMax;
    const float m0 = S.posLin[ea.x].w, m1 = S.posLin[ea.y].w, m2 = S.posLin[eb.x].w, m3 = S.posLin[eb.y].w;
    if (m0 > 0.0f) minMass = fminf(minMass, m0);
    if (m1 > 0.0f) minMass = fminf(minMass, m1);
    if (m2 > 0.0f) minMass = fminf(minMass, m2);
    if (m3 > 0.0f) minMass = fminf(minMass, m3);
    if (minMass == kFltMax) continue;
    if (EMIT) {
      const uint32_t slot = base + n;
      if (slot < P.maxSoft) {
        const V3 nGeo = d / dist;
        const uint32_t keyA = (kSoftEE << 28) | e, keyB = o;
        V3 nrm = nGeo;
        float g = dist;
        const int prev = soft_map_find(S, P, keyA, keyB);
        if (prev >= 0 && dot(nGeo, xyz(S.prevSoft[prev].normal)) < 0.0f) {
          nrm = -nGeo;
          g = -dist;
        }
        const float friction = sqrtf(fmaxf(
            (S.props[ea.x].w + S.props[ea.y].w) * 0.5f * (S.props[eb.x].w + S.props[eb.y].w) * 0.5f, 0.0f));
        S.soft[slot] = soft_record(S, P, U4{ea.x, ea.y, eb.x, eb.y}, F4{1.0f - s, s, -(1.0f - t), -t}, nrm, friction,
                                   v3s(0), g - hSum + P.elemMargin, hSum - P.elemMargin, soft_lambda_cap(P, minMass),
                                   minMass, kSoftEE, false, false, false, keyA, keyB);
      }
    }
    n++;
  }
  return n;
}

}  // namespace

// Per frame, before the trackers: the previous frame's tracked sets, the
// frame number the staggered reseeds follow (the frame index before this
// step), and cleared bounds.
__global__ void __launch_bounds__(256) k_iso_begin(DeviceState S, StepParams P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) S.isoFrame[0] += 1u;
  if (i < P.numBodies) {
    S.vtPrev[i] = S.vtTrack[i];
    S.ogcBound[i] = kOgcFarBits;
  }
  if (i < P.numContactEdges) S.eePrev[i] = S.eeTrack[i];
}

// element grid at the contact reach: each triangle / authored edge in every
// distinct bucket of its AABB padded by its largest vertex radius and the
// margin, at most 3 cells per axis (25_cloth.metal:147-267, isotropic
// branch); a non-finite element registers nowhere
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_iso_bin(DeviceState S, StepParams P) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numTris + P.numContactEdges) return;
  V3 lo, hi;
  float pad;
  uint32_t entry, key = kNone;
  if (gid < P.numTris) {
    const U4 t = S.tris[gid];
    const V3 a = P3(S, t.x), b = P3(S, t.y), c = P3(S, t.z);
    pad = fmaxf(fabsf(S.shape[t.x].w), fmaxf(fabsf(S.shape[t.y].w), fabsf(S.shape[t.z].w))) + P.elemMargin;
    lo = vmin(vmin(a, b), c);
    hi = vmax(vmax(a, b), c);
    entry = gid;
    if (!SCATTER) key = solid_key3(S, t.x, t.y, t.z);
  } else {
    const U2 e = S.edges[gid - P.numTris];
    const V3 a = P3(S, e.x), b = P3(S, e.y);
    pad = contact_pad(S, P, e.x, e.y);
    lo = vmin(a, b);
    hi = vmax(a, b);
    entry = kElemEdgeBit | (gid - P.numTris);
    if (!SCATTER) key = solid_key2(S, e.x, e.y);
  }
  lo = lo - v3s(pad);
  hi = hi + v3s(pad);
  if (!finite3(lo) || !finite3(hi)) return;
  const float cs = P.surfaceContactCellSize;
  const Cell3 c0 = cell_of(lo, cs), raw = cell_of(hi, cs);
  auto span = [](long long d) { return d < 0 ? 0 : (d > 2 ? 2 : (int)d); };
  const int nx = span((long long)raw.x - c0.x) + 1, ny = span((long long)raw.y - c0.y) + 1,
            nz = span((long long)raw.z - c0.z) + 1;
  const int n = nx * ny * nz;  // <= 27
  uint64_t seen = 0ull;  // one bit per registered bucket (mod 64), as in the element grid
  for (int i = 0; i < n; i++) {
    const int x = cell_add(c0.x, i % nx), y = cell_add(c0.y, (i / nx) % ny), z = cell_add(c0.z, i / (nx * ny));
    const uint32_t h = cell_hash(x, y, z, 0u, P.elemHashSize);
    const uint64_t bit = 1ull << (h & 63u);
    bool dup = false;
    if ((seen & bit) != 0ull)
      for (int j = 0; j < i && !dup; j++) {
        const int xj = cell_add(c0.x, j % nx), yj = cell_add(c0.y, (j / nx) % ny), zj = cell_add(c0.z, j / (nx * ny));
        dup = cell_hash(xj, yj, zj, 0u, P.elemHashSize) == h;
      }
    seen |= bit;
    if (dup) continue;
    const uint32_t slot = 2u * h + (gid < P.numTris ? 0u : 1u);
    if (SCATTER) S.elemCells[atomicAdd(&S.elemCellCursor[slot], 1u)] = entry;
    else {
      atomicAdd(&S.elemCellCount[slot], 1u);
      atomicMin(&S.isoSlotLo[slot], key);
      atomicMax(&S.isoSlotHi[slot], key);
    }
  }
}

// V-T trackers of pass 1 on 16-lane rows, one surface particle per row: the
// candidate pool (tracked set, triangles adjacent to its best, the mesh
// neighbours' best) one candidate per lane, and the 27-cell reseed with the
// lanes striding over each bucket.  Each lane keeps its own best four; a
// butterfly merge keeps the four smallest distinct (d2, id) of the union,
// which is what the serial scan keeps in any candidate order (an element's
// distance does not depend on who evaluates it), and the bound updates are
// minima, so the result equals the one-thread tracker's.
__device__ __forceinline__ void b4_add(Best4& b, uint32_t t, float d2) {
  if (t != kNone && !b4_has(b, t)) b4_insert(b, t, d2);
}
__device__ __forceinline__ void b4_row_merge(Best4& b) {
  for (int o = 8; o > 0; o >>= 1) {
    const uint32_t i0 = __shfl_xor(b.id0, o, 16), i1 = __shfl_xor(b.id1, o, 16);
    const uint32_t i2 = __shfl_xor(b.id2, o, 16), i3 = __shfl_xor(b.id3, o, 16);
    const float e0 = __shfl_xor(b.d0, o, 16), e1 = __shfl_xor(b.d1, o, 16);
    const float e2 = __shfl_xor(b.d2, o, 16), e3 = __shfl_xor(b.d3, o, 16);
    b4_add(b, i0, e0);
    b4_add(b, i1, e1);
    b4_add(b, i2, e2);
    b4_add(b, i3, e3);
  }
}

__global__ void __launch_bounds__(256) k_iso_track_vt(DeviceState S, StepParams P) {
  const uint32_t lane = threadIdx.x & 15u;
  const uint32_t p = blockIdx.x * 16u + (threadIdx.x >> 4);
  if (p == 0u && lane == 0u) S.softFlag[P.numParticles + P.numContactEdges] = 0u;
  if (p >= P.numParticles) return;  // whole row
  const uint32_t frame = S.isoFrame[0] - 1u;
  const float cs = P.surfaceContactCellSize;
  Best4 best;
  b4_init(best);
  float own = as_float(kOgcFarBits);
  VtQuery q;
  q.v = S.particleIdx[p];
  q.p = P3(S, q.v);
  q.gv = S.clothGroup[q.v];
  q.solid = q.gv != 0u && S.clothVert[q.v] == 0u;
  q.ns = S.nbrStart[q.v];
  q.ne = q.ns + S.nbrCount[q.v];
  q.n2s = S.nbr2Start[q.v];
  q.n2e = q.n2s + S.nbr2Count[q.v];
  q.cap2 = cs * cs;
  const U4 tr = S.vtPrev[q.v];
  // the pool: 4 tracked, 3 adjacent to the tracked best, 8 neighbours' best
  uint32_t t = kNone;
  if (lane < 4u) {
    t = u4_at(tr, (int)lane);
  } else if (lane < 7u) {
    if (tr.x < P.numTris) t = u4_at(S.triAdj[tr.x], (int)lane - 4);
  } else if (lane < 15u) {
    const uint32_t k = q.ns + (lane - 7u);
    if (k < q.ne) t = S.vtPrev[S.nbrList[k]].x;
  }
  vt_consider(S, P, q, best, own, t);
  b4_row_merge(best);
  V3 relV = xyz(S.velLin[q.v]);
  if (best.id0 < P.numTris) {
    const U4 bt = S.tris[best.id0];
    relV = relV - (xyz(S.velLin[bt.x]) + xyz(S.velLin[bt.y]) + xyz(S.velLin[bt.z])) / 3.0f;
  }
  if (iso_reseed(P, q.v, frame, relV, best.id0 == kNone, q.solid)) {
    const Cell3 cc = cell_of(q.p, cs);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const uint32_t h = cell_hash(cell_add(cc.x, dx), cell_add(cc.y, dy), cell_add(cc.z, dz), 0u,
                                       P.elemHashSize);
          if (own_solid_bucket(S, 2u * h, q.solid, q.gv)) continue;  // only candidates the solid filter rejects
          const uint32_t s0 = S.elemCellStart[2u * h], e0 = S.elemCellStart[2u * h + 1u];
          for (uint32_t k = s0 + lane; k < e0; k += 16u) vt_consider(S, P, q, best, own, S.elemCells[k]);
        }
    b4_row_merge(best);
  }
  for (int o = 8; o > 0; o >>= 1) own = fminf(own, __shfl_xor(own, o, 16));
  if (lane != 0u) return;
  const U4 kept = U4{best.id0, best.id1, best.id2, best.id3};
  S.vtTrack[q.v] = kept;
  if (own < as_float(kOgcFarBits)) atomicMin(&S.ogcBound[q.v], as_uint(own));
  S.softFlag[p] = vt_emit<false>(S, P, q.v, kept, 0u);
}

// Trackers (pass 1: tracked sets, bounds, contact counts into softFlag) and
// emission (pass 2: records from the tracked sets at the rigid-feature count
// plus softStart).  One thread per producer.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_iso_track(DeviceState S, StepParams P, uint32_t first) {
  const uint32_t p = first + blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nProd = P.numParticles + P.numContactEdges;
  if (p >= nProd) return;
  if (EMIT) {
    const uint32_t base = S.counters[CTR_SOFT_RT] + S.softStart[p];
    if (p < P.numParticles) {
      const uint32_t v = S.particleIdx[p];
      vt_emit<true>(S, P, v, S.vtTrack[v], base);
    } else {
      const uint32_t e = p - P.numParticles;
      ee_emit<true>(S, P, e, S.eeTrack[e], base);
    }
    if (p == nProd - 1u) S.counters[CTR_SOFT_ISO] = S.softStart[nProd];
    return;
  }
  const uint32_t frame = S.isoFrame[0] - 1u;
  const float cs = P.surfaceContactCellSize;
  Best4 best;
  b4_init(best);
  float own = as_float(kOgcFarBits);
  if (p < P.numParticles) {
    VtQuery q;
    q.v = S.particleIdx[p];
    q.p = P3(S, q.v);
    q.gv = S.clothGroup[q.v];
    q.solid = q.gv != 0u && S.clothVert[q.v] == 0u;
    q.ns = S.nbrStart[q.v];
    q.ne = q.ns + S.nbrCount[q.v];
    q.n2s = S.nbr2Start[q.v];
    q.n2e = q.n2s + S.nbr2Count[q.v];
    q.cap2 = cs * cs;
    const U4 tr = S.vtPrev[q.v];
    vt_consider(S, P, q, best, own, tr.x);
    vt_consider(S, P, q, best, own, tr.y);
    vt_consider(S, P, q, best, own, tr.z);
    vt_consider(S, P, q, best, own, tr.w);
    if (tr.x < P.numTris) {
      const U4 adj = S.triAdj[tr.x];
      vt_consider(S, P, q, best, own, adj.x);
      vt_consider(S, P, q, best, own, adj.y);
      vt_consider(S, P, q, best, own, adj.z);
    }
    for (uint32_t k = q.ns; k < q.ne && k < q.ns + 8u; k++) vt_consider(S, P, q, best, own, S.vtPrev[S.nbrList[k]].x);
    V3 relV = xyz(S.velLin[q.v]);
    if (best.id0 < P.numTris) {
      const U4 bt = S.tris[best.id0];
      relV = relV - (xyz(S.velLin[bt.x]) + xyz(S.velLin[bt.y]) + xyz(S.velLin[bt.z])) / 3.0f;
    }
    if (iso_reseed(P, q.v, frame, relV, best.id0 == kNone, q.solid)) {
      const Cell3 cc = cell_of(q.p, cs);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            const uint32_t h = cell_hash(cell_add(cc.x, dx), cell_add(cc.y, dy), cell_add(cc.z, dz), 0u,
                                         P.elemHashSize);
            if (own_solid_bucket(S, 2u * h, q.solid, q.gv)) continue;  // only candidates the solid filter rejects
            const uint32_t s0 = S.elemCellStart[2u * h], e0 = S.elemCellStart[2u * h + 1u];
            for (uint32_t k = s0; k < e0; k++) vt_consider(S, P, q, best, own, S.elemCells[k]);
          }
    }
    const U4 kept = U4{best.id0, best.id1, best.id2, best.id3};
    S.vtTrack[q.v] = kept;
    if (own < as_float(kOgcFarBits)) atomicMin(&S.ogcBound[q.v], as_uint(own));
    S.softFlag[p] = vt_emit<false>(S, P, q.v, kept, 0u);
    return;
  }
  EeQuery q;
  q.e = p - P.numParticles;
  q.a = S.edges[q.e];
  q.a0 = P3(S, q.a.x);
  q.a1 = P3(S, q.a.y);
  q.ga = S.clothGroup[q.a.x];
  q.solid = q.ga != 0u && S.clothGroup[q.a.y] == q.ga && S.clothVert[q.a.x] == 0u && S.clothVert[q.a.y] == 0u;
  q.ns0 = S.nbrStart[q.a.x];
  q.ne0 = q.ns0 + S.nbrCount[q.a.x];
  q.ns1 = S.nbrStart[q.a.y];
  q.ne1 = q.ns1 + S.nbrCount[q.a.y];
  q.n2s = S.nbr2Start[q.a.x];
  q.n2e = q.n2s + S.nbr2Count[q.a.x];
  q.cap2 = cs * cs;
  const U4 tr = S.eePrev[q.e];
  ee_consider(S, P, q, best, own, tr.x);
  ee_consider(S, P, q, best, own, tr.y);
  ee_consider(S, P, q, best, own, tr.z);
  ee_consider(S, P, q, best, own, tr.w);
  for (int side = 0; side < 2; side++) {
    const uint32_t vv = side == 0 ? q.a.x : q.a.y;
    const uint32_t es = S.vEdgeStart[vv], ec = min(S.vEdgeCount[vv], 6u);
    for (uint32_t k = es; k < es + ec; k++) {
      const uint32_t ie = S.vEdgeList[k];
      if (ie < P.numContactEdges) ee_consider(S, P, q, best, own, S.eePrev[ie].x);
    }
  }
  const V3 velA = (xyz(S.velLin[q.a.x]) + xyz(S.velLin[q.a.y])) * 0.5f;
  V3 relV = velA;
  if (best.id0 < P.numContactEdges) {
    const U2 bt = S.edges[best.id0];
    relV = relV - (xyz(S.velLin[bt.x]) + xyz(S.velLin[bt.y])) * 0.5f;
  }
  if (iso_reseed(P, q.e, frame, relV, best.id0 == kNone, q.solid)) {
    const Cell3 cc = cell_of((q.a0 + q.a1) * 0.5f, cs);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const uint32_t h =
              cell_hash(cell_add(cc.x, dx), cell_add(cc.y, dy), cell_add(cc.z, dz), 0u, P.elemHashSize);
          if (own_solid_bucket(S, 2u * h + 1u, q.solid, q.ga)) continue;
          const uint32_t s0 = S.elemCellStart[2u * h + 1u], e0 = S.elemCellStart[2u * h + 2u];
          for (uint32_t k = s0; k < e0; k++) ee_consider(S, P, q, best, own, S.elemCells[k] & ~kElemEdgeBit);
        }
  }
  const U4 kept = U4{best.id0, best.id1, best.id2, best.id3};
  S.eeTrack[q.e] = kept;
  if (own < as_float(kOgcFarBits)) {
    atomicMin(&S.ogcBound[q.a.x], as_uint(own));
    atomicMin(&S.ogcBound[q.a.y], as_uint(own));
  }
  S.softFlag[p] = ee_emit<false>(S, P, q.e, kept, 0u);
}

// Conditional in-iteration refresh (25_cloth.metal:857-939): when more than
// 1 % of the surface particles were held back by their bound since the last
// check, each re-anchors at its live pose with its bound recomputed from the
// tracked sets.  Every workgroup reads the truncation count, then counts
// itself in; the last one to arrive clears both counters for the next
// iteration.
__global__ void __launch_bounds__(256) k_ogc_refresh(DeviceState S, StepParams P) {
  if (P.convCheck && S.counters[CTR_CONV_DONE] != 0u) return;  // the frame's residual check has ended its iterations
  __shared__ uint32_t fire;
  if (threadIdx.x == 0) {
    const uint32_t exceeded = __hip_atomic_load(&S.counters[CTR_OGC], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    fire = (uint64_t)exceeded * 100u > (uint64_t)P.numParticles ? 1u : 0u;
    __threadfence();
    if (atomicAdd(&S.counters[CTR_OGC_ARRIVE], 1u) == gridDim.x - 1u) {
      S.counters[CTR_OGC] = 0u;
      S.counters[CTR_OGC_ARRIVE] = 0u;
      if (fire) S.counters[CTR_OGC_REFRESH] += 1u;
    }
  }
  __syncthreads();
  if (!fire) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.numParticles) return;
  const uint32_t v = S.particleIdx[i];
  const V3 pv = P3(S, v);
  const uint32_t n2s = S.nbr2Start[v], n2e = n2s + S.nbr2Count[v];
  float dmin = 3.0e38f;
  const U4 tr = S.vtTrack[v];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t t = u4_at(tr, k);
    if (t >= P.numTris) break;
    const U4 tv = S.tris[t];
    if (nbr_contains(S.nbr2List, n2s, n2e, tv.x) || nbr_contains(S.nbr2List, n2s, n2e, tv.y) ||
        nbr_contains(S.nbr2List, n2s, n2e, tv.z))
      continue;
    V3 bary;
    const V3 d = pv - closest_pt_triangle(pv, P3(S, tv.x), P3(S, tv.y), P3(S, tv.z), bary);
    dmin = fminf(dmin, dot(d, d));
  }
  const uint32_t es = S.vEdgeStart[v], ec = min(S.vEdgeCount[v], 6u);
  for (uint32_t k = es; k < es + ec; k++) {
    const uint32_t me = S.vEdgeList[k];
    if (me >= P.numContactEdges) continue;
    const uint32_t other = S.eeTrack[me].x;
    if (other >= P.numContactEdges) continue;
    const U2 eo = S.edges[other];
    if (nbr_contains(S.nbr2List, n2s, n2e, eo.x) || nbr_contains(S.nbr2List, n2s, n2e, eo.y)) continue;
    const U2 em = S.edges[me];
    const V3 a0 = P3(S, em.x), a1 = P3(S, em.y), b0 = P3(S, eo.x), b1 = P3(S, eo.y);
    float s, t;
    closest_seg_seg(a0, a1, b0, b1, s, t);
    const V3 d = (a0 + (a1 - a0) * s) - (b0 + (b1 - b0) * t);
    dmin = fminf(dmin, dot(d, d));
  }
  S.ogcBound[v] = as_uint(dmin);
  S.ogcPrev[v] = S.posLin[v];
}

// ---------------------------------------------------------------------------
// 1-ring velocity viscosity of surface particles (40_solver.metal:2154-2177)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_visc_copy(DeviceState S, StepParams P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.numBodies) S.velScratch[i] = S.velLin[i];
}
__global__ void __launch_bounds__(256) k_viscosity(DeviceState S, StepParams P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.numParticles) return;
  uint32_t v = S.particleIdx[i];
  uint32_t s = S.nbrStart[v], n = S.nbrCount[v];
  if (n == 0) return;
  V3 avg = v3s(0);
  for (uint32_t k = s; k < s + n; k++) avg += xyz(S.velScratch[S.nbrList[k]]);
  avg = avg / (float)n;
  V3 vel = xyz(S.velScratch[v]);
  S.velLin[v] = f4(vel + (avg - vel) * P.clothViscosity, 0.0f);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static uint32_t cap_blocks(uint32_t n, uint32_t cap) {
  uint32_t b = (n + 255) / 256;
  if (b > cap) b = cap;
  return b ? b : 1;
}

void launch_rt_emit(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (P.numTris == 0) return;
  uint32_t blocks = (P.numTris + 255) / 256;
  hipLaunchKernelGGL(k_rt_emit<false>, dim3(blocks), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.softFlag, S.softStart, P.numTris + 1, S.scanScratch);
  hipLaunchKernelGGL(k_rt_emit<true>, dim3(blocks), dim3(256), 0, st, S, P);
}

void launch_dat_pairs(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  if (!P.pairStream) return;
  uint32_t nElem = P.numTris + P.numEdges;
  uint32_t eb = (nElem + 255) / 256;
  const uint32_t slots = 2u * P.elemHashSize;
  launch_fill_u32(st, S.elemCellCount, slots + 1u, 0u);
  hipLaunchKernelGGL(k_el_bin<false>, dim3(eb), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.elemCellCount, S.elemCellStart, slots + 1u, S.scanScratch);
  hipLaunchKernelGGL(k_copy_u32, dim3(cap_blocks(slots, gridCap)), dim3(256), 0, st, S.elemCellCursor,
                     S.elemCellStart, slots);
  hipLaunchKernelGGL(k_el_bin<true>, dim3(eb), dim3(256), 0, st, S, P);
  uint32_t nProd = P.numParticles + P.numEdges;
  uint32_t pb = (nProd + 3) / 4;
  hipLaunchKernelGGL(k_dat_count, dim3((nProd + 15) / 16), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.datCount, S.datStart, nProd + 1, S.scanScratch);
  hipLaunchKernelGGL(k_dat_pairs, dim3(pb), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_dat_finalize, dim3(1), dim3(1), 0, st, S, P);
}

// the pair stream's contacts after the rigid-feature ones (without the
// stream, the rigid-feature contacts alone), then the persistence map
void launch_dat_contacts(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  if (P.numTris == 0) return;
  if (P.pairStream) {
    uint32_t blocks = cap_blocks(P.maxDatPairs, gridCap);
    hipLaunchKernelGGL(k_dat_contacts<false>, dim3(blocks), dim3(256), 0, st, S, P);
    launch_exclusive_scan_dev(st, S.softFlag, S.softStart, P.maxDatPairs + 1,
                              S.datStart + P.numParticles + P.numEdges, P.maxDatPairs, S.scanScratch);
    hipLaunchKernelGGL(k_dat_contacts<true>, dim3(blocks), dim3(256), 0, st, S, P);
  }
  hipLaunchKernelGGL(k_soft_finalize, dim3(1), dim3(1), 0, st, S, P);
  hipLaunchKernelGGL(k_softmap_clear, dim3(cap_blocks(P.softMapCapacity, gridCap)), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_softmap_insert, dim3(cap_blocks(P.maxSoft, gridCap)), dim3(256), 0, st, S, P);
}

void launch_dat_predictor_pass(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  if (!P.datOn) return;
  hipLaunchKernelGGL(k_dat_reduce, dim3(cap_blocks(P.maxDatPairs, gridCap)), dim3(256), 0, st, S, P,
                     (const F4*)S.initLin, kDatFullPass);
  if (P.numParticles)
    hipLaunchKernelGGL(k_dat_apply, dim3((P.numParticles + 255) / 256), dim3(256), 0, st, S, P,
                       (const F4*)S.initLin, 1u, 0u, 0u);
}

void launch_dat_reanchor(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (!P.datOn || P.numParticles == 0) return;
  hipLaunchKernelGGL(k_dat_reanchor, dim3((P.numParticles + 255) / 256), dim3(256), 0, st, S, P);
}

void launch_dat_color_pass(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t color, uint32_t start,
                           uint32_t count, uint32_t gridCap) {
  if (!P.datOn || count == 0) return;
  hipLaunchKernelGGL(k_dat_reduce, dim3(cap_blocks(P.maxDatPairs, gridCap)), dim3(256), 0, st, S, P,
                     (const F4*)S.datAnchor, color);
  hipLaunchKernelGGL(k_dat_apply, dim3((count + 255) / 256), dim3(256), 0, st, S, P, (const F4*)S.datAnchor, 0u,
                     start, count);
}

void launch_dat_restore(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (!P.datOn) return;
  hipLaunchKernelGGL(k_dat_restore, dim3((P.numBodies + 255) / 256), dim3(256), 0, st, S, P);
}

// the isotropic mode's per-frame work at the step-start pose, after the
// rigid-feature records: tracked sets, bounds and the tracked contacts
void launch_iso_track(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  if (!P.isoOn) return;
  const uint32_t nBegin = P.numBodies > P.numContactEdges ? P.numBodies : P.numContactEdges;
  hipLaunchKernelGGL(k_iso_begin, dim3((nBegin + 255) / 256), dim3(256), 0, st, S, P);
  const uint32_t nElem = P.numTris + P.numContactEdges, slots = 2u * P.elemHashSize;
  launch_fill_u32(st, S.elemCellCount, slots + 1u, 0u);
  launch_fill_u32(st, S.isoSlotLo, slots, kNone);
  launch_fill_u32(st, S.isoSlotHi, slots, 0u);
  hipLaunchKernelGGL(k_iso_bin<false>, dim3((nElem + 255) / 256), dim3(256), 0, st, S, P);
  launch_exclusive_scan(st, S.elemCellCount, S.elemCellStart, slots + 1u, S.scanScratch);
  hipLaunchKernelGGL(k_copy_u32, dim3(cap_blocks(slots, gridCap)), dim3(256), 0, st, S.elemCellCursor,
                     S.elemCellStart, slots);
  hipLaunchKernelGGL(k_iso_bin<true>, dim3((nElem + 255) / 256), dim3(256), 0, st, S, P);
  const uint32_t nProd = P.numParticles + P.numContactEdges;
  // pass 1: V-T trackers on 16-lane rows (also clearing the scan's total
  // slot), E-E trackers one thread each
  hipLaunchKernelGGL(k_iso_track_vt, dim3((P.numParticles + 15) / 16 > 0 ? (P.numParticles + 15) / 16 : 1), dim3(256), 0,
                     st, S, P);
  if (P.numContactEdges > 0)
    hipLaunchKernelGGL(k_iso_track<false>, dim3((P.numContactEdges + 255) / 256), dim3(256), 0, st, S, P,
                       P.numParticles);
  launch_exclusive_scan(st, S.softFlag, S.softStart, nProd + 1u, S.scanScratch);
  hipLaunchKernelGGL(k_iso_track<true>, dim3((nProd + 255) / 256), dim3(256), 0, st, S, P, 0u);
}

void launch_ogc_refresh(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (!P.isoOn || P.numParticles == 0) return;
  hipLaunchKernelGGL(k_ogc_refresh, dim3((P.numParticles + 255) / 256), dim3(256), 0, st, S, P);
}

void launch_viscosity(hipStream_t st, const DeviceState& S, const StepParams& P) {
  if (P.numTris == 0 || P.numParticles == 0 || !(P.clothViscosity > 0.0f)) return;
  hipLaunchKernelGGL(k_visc_copy, dim3((P.numBodies + 255) / 256), dim3(256), 0, st, S, P);
  hipLaunchKernelGGL(k_viscosity, dim3((P.numParticles + 255) / 256), dim3(256), 0, st, S, P);
}

void launch_soft_map_find(hipStream_t st, const DeviceState& S, const StepParams& P, const U2* q, int32_t* out,
                          uint32_t n) {
  if (n == 0 || P.softMapCapacity == 0) return;
  hipLaunchKernelGGL(k_softmap_query, dim3((n + 255) / 256), dim3(256), 0, st, S, P, q, out, n);
}

}  // namespace avbd
