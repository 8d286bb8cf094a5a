// narrow.hip -- collider-pair narrow phase for gfx950
// (semantics: 30_narrowphase.metal:189-247, 337-1322; warm start :1265-1321).
//
// One 16-lane row of a wave64 owns one collider pair; a 256-thread workgroup
// holds 16 rows and the grid strides over this frame's pair count.  Work is
// spread over the row's lanes instead of running one serial routine per pair:
//   * box-box: lanes 0-14 each test one separating-axis candidate (faces of
//     A, faces of B, the nine edge cross products).  One row ballot says
//     whether any axis separates; two butterfly arg-max reductions over
//     (depth, axis index) pick the deepest face and edge axes, lowest axis on
//     ties -- the serial first-wins order;
//   * the incident face is clipped against the reference face's four side
//     planes with one polygon vertex per lane.  Each lane emits the crossing
//     point of the edge entering its vertex and then its vertex if inside; two
//     row ballots give every emission its place (prefix count), and the new
//     polygon is read back from LDS.  The 16-vertex polygon bound is kept;
//   * each surviving vertex is a candidate contact on its own lane.  The 1 mm
//     merge of near duplicates is a pairwise mask per lane followed by an
//     in-order acceptance pass of 16 row ballots, so the kept contacts and
//     their order are those of a serial greedy scan;
//   * capsule-box seeds run one per lane; convex-hull vertices are evaluated
//     16 at a time and the selection rules (tolerance ties) are applied in
//     vertex order to row-broadcast values;
//   * each kept contact's warm start (feature key or proximity match against
//     the previous frame's manifold of the same collider pair) and its record
//     write run on the contact's lane; lane 0 writes the manifold header.
//
// Contact counts and feature keys must equal the oracle's, so contraction is
// off: every geometric quantity rounds as written.
#include "kernels.hpp"
#include "pairmap.hpp"

#pragma clang fp contract(off)

namespace avbd {

namespace {

constexpr int kRow = 16;                  // lanes per collider pair
constexpr int kRowsPerGroup = 16;         // 256-thread workgroups
constexpr int kPolyMax = 16;              // clipped polygon bound
constexpr float kDegenerate = 1.0e-6f;    // squared axis length / denominators treated as zero
constexpr float kSlop = 1.0e-5f;          // clip-plane and depth tolerance
constexpr float kMergeSq = 1.0e-6f;       // contacts whose midpoints lie within 1 mm merge
constexpr float kHuge = 3.402823466e38f;

// ---- row primitives (a row = 16 consecutive lanes of the wave) ----
__device__ __forceinline__ int row_lane() { return (int)(threadIdx.x & 15u); }
// bit k = predicate of the row's lane k
__device__ __forceinline__ uint32_t row_bits(bool p) {
  return (uint32_t)(__ballot(p) >> (threadIdx.x & 48u)) & 0xFFFFu;
}
__device__ __forceinline__ float row_get(float v, int k) { return __shfl(v, k, kRow); }
__device__ __forceinline__ int row_get(int v, int k) { return __shfl(v, k, kRow); }
__device__ __forceinline__ uint32_t row_get(uint32_t v, int k) { return __shfl(v, k, kRow); }
__device__ __forceinline__ V3 row_get(V3 v, int k) { return v3(row_get(v.x, k), row_get(v.y, k), row_get(v.z, k)); }

// LDS stores of a lane visible to the other lanes of its wave
__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row-wide arg-max of (ok, depth); the lowest axis index wins ties.  The
// result is taken from lane 0 so every lane holds the same answer.
__device__ __forceinline__ void row_deepest(bool& ok, float& depth, int& axis) {
  for (int o = 8; o > 0; o >>= 1) {
    const bool ok2 = __shfl_xor((int)ok, o, kRow) != 0;
    const float d2 = __shfl_xor(depth, o, kRow);
    const int a2 = __shfl_xor(axis, o, kRow);
    if (ok2 && (!ok || d2 > depth || (d2 == depth && a2 < axis))) {
      ok = true;
      depth = d2;
      axis = a2;
    }
  }
  ok = row_get((int)ok, 0) != 0;
  depth = row_get(depth, 0);
  axis = row_get(axis, 0);
}

// component-wise selects: a ternary on an aggregate can become a select of
// two stack addresses, which puts both operands in scratch memory
__device__ __forceinline__ V3 pick(bool c, V3 a, V3 b) { return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
__device__ __forceinline__ F4 pick(bool c, F4 a, F4 b) {
  return F4{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w};
}

// ---- geometry of one collider at the current pose ----
struct Obb {
  V3 c, e;          // center, half extents
  V3 u0, u1, u2;    // local axes in world space
};
__device__ __forceinline__ V3 obb_u(const Obb& b, int k) { return pick(k == 0, b.u0, pick(k == 1, b.u1, b.u2)); }
__device__ __forceinline__ Obb pick(bool c, const Obb& a, const Obb& b) {
  Obb o;
  o.c = pick(c, a.c, b.c);
  o.e = pick(c, a.e, b.e);
  o.u0 = pick(c, a.u0, b.u0);
  o.u1 = pick(c, a.u1, b.u1);
  o.u2 = pick(c, a.u2, b.u2);
  return o;
}
__device__ __forceinline__ Obb make_obb(V3 c, F4 q, V3 size) {
  Obb b;
  b.c = c;
  b.e = size * 0.5f;
  b.u0 = q_rotate(q, v3(1, 0, 0));
  b.u1 = q_rotate(q, v3(0, 1, 0));
  b.u2 = q_rotate(q, v3(0, 0, 1));
  return b;
}
// half-width of the box's projection on the unit direction n
__device__ __forceinline__ float obb_reach(const Obb& b, V3 n) {
  return b.e.x * fabsf(dot(n, b.u0)) + b.e.y * fabsf(dot(n, b.u1)) + b.e.z * fabsf(dot(n, b.u2));
}
// the vertex furthest along d
__device__ __forceinline__ V3 obb_extreme(const Obb& b, V3 d) {
  const float sx = dot(d, b.u0) >= 0.0f ? 1.0f : -1.0f;
  const float sy = dot(d, b.u1) >= 0.0f ? 1.0f : -1.0f;
  const float sz = dot(d, b.u2) >= 0.0f ? 1.0f : -1.0f;
  return b.c + b.u0 * (b.e.x * sx) + b.u1 * (b.e.y * sy) + b.u2 * (b.e.z * sz);
}
// the edge parallel to axis k that lies furthest along d
__device__ __forceinline__ void obb_extreme_edge(const Obb& b, int k, V3 d, V3& p0, V3& p1) {
  const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
  const float s1 = dot(d, obb_u(b, k1)) >= 0.0f ? 1.0f : -1.0f;
  const float s2 = dot(d, obb_u(b, k2)) >= 0.0f ? 1.0f : -1.0f;
  const V3 mid = b.c + obb_u(b, k1) * (comp(b.e, k1) * s1) + obb_u(b, k2) * (comp(b.e, k2) * s2);
  p0 = mid - obb_u(b, k) * comp(b.e, k);
  p1 = mid + obb_u(b, k) * comp(b.e, k);
}
// in-plane axes and half extents of the face whose normal is axis k
__device__ __forceinline__ void obb_face(const Obb& b, int k, V3& u, V3& v, float& eu, float& ev) {
  u = pick(k == 0, b.u1, b.u0);
  eu = k == 0 ? b.e.y : b.e.x;
  v = pick(k == 2, b.u1, b.u2);
  ev = k == 2 ? b.e.y : b.e.z;
}

__device__ __forceinline__ V3 clamp_box(V3 p, V3 h) { return vmin(vmax(p, -h), h); }
__device__ __forceinline__ V3 nearest_on_segment(V3 p, V3 a, V3 b) {
  const V3 ab = b - a;
  const float t = clampf(dot(p - a, ab) / fmaxf(dot(ab, ab), 1e-12f), 0.0f, 1.0f);
  return a + ab * t;
}
// closest points of segments [p0, p1] and [q0, q1] (clamped parameters)
__device__ __forceinline__ void nearest_between_segments(V3 p0, V3 p1, V3 q0, V3 q1, V3& cp, V3& cq) {
  const V3 dp = p1 - p0, dq = q1 - q0, w = p0 - q0;
  const float pp = dot(dp, dp), qq = dot(dq, dq), qw = dot(dq, w);
  float s = 0.0f, t = 0.0f;
  if (pp <= kDegenerate && qq <= kDegenerate) {
    cp = p0;
    cq = q0;
    return;
  }
  if (pp <= kDegenerate) {
    t = clampf(qw / qq, 0.0f, 1.0f);
  } else {
    const float pw = dot(dp, w);
    if (qq <= kDegenerate) {
      s = clampf(-pw / pp, 0.0f, 1.0f);
    } else {
      const float pq = dot(dp, dq);
      const float den = pp * qq - pq * pq;
      if (fabsf(den) > kDegenerate) s = clampf((pq * qw - pw * qq) / den, 0.0f, 1.0f);
      t = (pq * s + qw) / qq;
      if (t < 0.0f) {
        t = 0.0f;
        s = clampf(-pw / pp, 0.0f, 1.0f);
      } else if (t > 1.0f) {
        t = 1.0f;
        s = clampf((pq - pw) / pp, 0.0f, 1.0f);
      }
    }
  }
  cp = p0 + dp * s;
  cq = q0 + dq * t;
}
// index of the smallest component (first on ties)
__device__ __forceinline__ int min_axis(V3 v) {
  int k = 0;
  if (v.y < comp(v, k)) k = 1;
  if (v.z < comp(v, k)) k = 2;
  return k;
}

// speculative contact distance: velocity-aware, capped by the smaller radius
__device__ __forceinline__ float spec_cap(float ra, float rb, float margin) {
  const float r = fminf(fabsf(ra), fabsf(rb));
  return fminf(0.25f, fmaxf(4.0f * margin, 3.0f * r));
}
__device__ __forceinline__ float detect_margin(V3 n, V3 rel, const StepParams& P, float ra, float rb) {
  const float closing = fmaxf(0.0f, -dot(rel, n));
  return P.collisionMargin + fminf(closing * P.dt, spec_cap(ra, rb, P.collisionMargin));
}
// contact penalty floor: lighter dynamic body's m / dt^2, clamped to the caps
__device__ __forceinline__ V3 contact_penalty_floor(const DeviceState& S, uint32_t ba, uint32_t bb,
                                                    const StepParams& P) {
  const float mA = S.posLin[ba].w, mB = S.posLin[bb].w;
  const float mMin = fminf(mA > 0.0f ? mA : kHuge, mB > 0.0f ? mB : kHuge);
  if (mMin == kHuge) return v3s(kPenaltyMin);
  const float k = fminf(kPenaltyMax, fmaxf(kPenaltyMin, mMin / fmaxf(P.dt * P.dt, 1e-12f)));
  return v3(k, fminf(k, kPenaltyMaxT), fminf(k, kPenaltyMaxT));
}

// One clip stage over the row's polygon (vertex k on lane k, n vertices):
// keeps the side dot(pn, x) - off <= kSlop.  Lane k owns the edge entering
// vertex k from vertex k-1 (cyclic).  Returns the new vertex count.
__device__ __forceinline__ int row_clip(V3& pv, int n, V3 pn, float off, F4* poly) {
  if (n <= 0) return 0;
  const int k = row_lane();
  const bool live = k < n;
  const float db = dot(pn, pv) - off;
  const int from = k == 0 ? n - 1 : k - 1;
  const float da = row_get(db, from);
  const V3 pa = row_get(pv, from);
  const bool aIn = da <= kSlop, bIn = db <= kSlop;
  const bool cuts = live && aIn != bIn;
  const bool keeps = live && bIn;
  const uint32_t below = (1u << k) - 1u;
  const uint32_t mc = row_bits(cuts), mk = row_bits(keeps);
  int at = __popc(mc & below) + __popc(mk & below);
  if (cuts) {
    float t = 0.0f;
    const float den = da - db;
    if (fabsf(den) > kDegenerate) t = clampf(da / den, 0.0f, 1.0f);
    if (at < kPolyMax) poly[at] = f4(pa + (pv - pa) * t, 0.0f);
    at++;
  }
  if (keeps && at < kPolyMax) poly[at] = f4(pv, 0.0f);
  lds_wave_sync();
  const int total = min(__popc(mc) + __popc(mk), kPolyMax);
  if (k < total) pv = xyz(poly[k]);
  lds_wave_sync();
  return total;
}

// per-lane contact candidate
struct Hit {
  bool on;          // this lane carries a kept contact
  int slot;         // its index in the manifold
  V3 xA, xB;        // world points on collider A and B
  uint32_t key;     // feature key
};

// ---- box vs box ----
// Returns false when the pair is separated; otherwise fills the lane's hit
// and the normal from B toward A.
__device__ __forceinline__ bool box_box(const Obb& A, const Obb& B, V3 rel, float cap, const StepParams& P, F4* lds, Hit& h,
                        V3& nrm) {
  const int k = row_lane();
  const V3 delta = B.c - A.c;
  V3 axis = v3s(0);
  if (k < 3) axis = obb_u(A, k);
  else if (k < 6) axis = obb_u(B, k - 3);
  else if (k < 15) axis = cross(obb_u(A, (k - 6) / 3), obb_u(B, (k - 6) % 3));
  const float l2 = dot(axis, axis);
  bool ok = k < 15 && !(l2 < kDegenerate);
  V3 n = v3s(0);
  float depth = -kHuge;
  bool apart = false;
  if (ok) {
    n = axis * rsq(l2);
    if (dot(n, delta) < 0.0f) n = -n;
    depth = fabsf(dot(delta, n)) - (obb_reach(A, n) + obb_reach(B, n));
    apart = depth > fminf(fmaxf(0.0f, -dot(rel, n)) * P.dt, cap);
  }
  if (row_bits(apart) != 0) return false;
  bool fOk = ok && k < 6, eOk = ok && k >= 6;
  float fDepth = depth, eDepth = depth;
  int fAxis = k, eAxis = k;
  row_deepest(fOk, fDepth, fAxis);
  row_deepest(eOk, eDepth, eAxis);
  if (!fOk) return false;
  int best = fAxis;
  float bestDepth = fDepth;
  if (eOk && 0.95f * eDepth > fDepth + 0.01f) {
    best = eAxis;
    bestDepth = eDepth;
  }
  const V3 bn = row_get(n, best);  // from A toward B
  nrm = -bn;
  h.on = false;
  h.slot = 0;
  if (best >= 6) {
    // edge-edge: one contact at the closest points of the two extreme edges
    const int ea = (best - 6) / 3, eb = (best - 6) % 3;
    V3 a0, a1, b0, b1;
    obb_extreme_edge(A, ea, bn, a0, a1);
    obb_extreme_edge(B, eb, -bn, b0, b1);
    nearest_between_segments(a0, a1, b0, b1, h.xA, h.xB);
    h.key = (2u << 24) | ((uint32_t)ea << 8) | (uint32_t)eb;
    h.on = k == 0;
    return true;
  }
  // face contact: clip the incident face to the reference face's side planes
  const bool refIsA = best < 3;
  const int rAxis = refIsA ? best : best - 3;
  const Obb ref = pick(refIsA, A, B);
  const Obb inc = pick(refIsA, B, A);
  const V3 outward = pick(refIsA, bn, -bn);
  const V3 rN = obb_u(ref, rAxis) * (dot(outward, obb_u(ref, rAxis)) >= 0.0f ? 1.0f : -1.0f);
  const V3 rC = ref.c + rN * comp(ref.e, rAxis);
  V3 rU, rV;
  float rEU, rEV;
  obb_face(ref, rAxis, rU, rV, rEU, rEV);
  int iAxis = 0;
  float iAlign = -kHuge;
  for (int j = 0; j < 3; j++) {
    const float a = fabsf(dot(obb_u(inc, j), rN));
    if (a > iAlign) {
      iAlign = a;
      iAxis = j;
    }
  }
  const V3 iN = obb_u(inc, iAxis) * (dot(obb_u(inc, iAxis), rN) > 0.0f ? -1.0f : 1.0f);
  const V3 iC = inc.c + iN * comp(inc.e, iAxis);
  V3 iU, iV;
  float iEU, iEV;
  obb_face(inc, iAxis, iU, iV, iEU, iEV);
  // incident face corners in the order (+u+v, -u+v, -u-v, +u-v), one per lane
  V3 pv = iC + iU * ((k == 0 || k == 3) ? iEU : -iEU) + iV * (k < 2 ? iEV : -iEV);
  int cnt = 4;
  cnt = row_clip(pv, cnt, rU, dot(rU, rC) + rEU, lds);
  cnt = row_clip(pv, cnt, -rU, dot(-rU, rC) + rEU, lds);
  cnt = row_clip(pv, cnt, rV, dot(rV, rC) + rEV, lds);
  cnt = row_clip(pv, cnt, -rV, dot(-rV, rC) + rEV, lds);
  const uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | ((uint32_t)rAxis << 16) | ((uint32_t)iAxis << 8);
  // vertices at most the penetration depth (plus slop) above the reference face
  const float slop = kSlop + fmaxf(0.0f, bestDepth);
  bool cand = false;
  V3 xA = v3s(0), xB = v3s(0);
  if (k < cnt) {
    const float d = dot(pv - rC, rN);
    if (!(d > slop)) {
      cand = true;
      const V3 onRef = pv - rN * d;
      xA = pick(refIsA, onRef, pv);
      xB = pick(refIsA, pv, onRef);
    }
  }
  // pairwise merge mask against the earlier candidates' midpoints
  const V3 mid = (xA + xB) * 0.5f;
  if (cand) lds[k] = f4(mid, 0.0f);
  const uint32_t candBits = row_bits(cand);
  lds_wave_sync();
  uint32_t close = 0;
  if (cand) {
    for (int j = 0; j < k; j++) {
      if (!((candBits >> j) & 1u)) continue;
      const V3 d = mid - xyz(lds[j]);
      if (dot(d, d) < kMergeSq) close |= 1u << j;
    }
  }
  lds_wave_sync();
  // in-order acceptance: a candidate is kept unless it merges with a kept one
  // or the manifold is full
  uint32_t kept = 0;
  for (int j = 0; j < kRow; j++) {
    const bool take = k == j && cand && (close & kept) == 0u && __popc(kept) < kMaxContacts;
    kept |= row_bits(take);
  }
  if (kept == 0u) {
    // nothing survived the depth test: one contact between the extreme vertices
    h.xA = obb_extreme(A, bn);
    h.xB = obb_extreme(B, -bn);
    h.key = prefix;
    h.on = k == 0;
    return true;
  }
  h.on = (kept >> k) & 1u;
  h.slot = __popc(kept & ((1u << k) - 1u));
  h.xA = xA;
  h.xB = xB;
  h.key = prefix | (uint32_t)k;
  return true;
}

// ---- convex hull vertex vs box ----
struct HullVtx {
  V3 xHull, xBox, n;  // hull vertex, box point, box -> hull normal
  float depth;        // signed separation (negative inside)
  uint32_t key;
};
__device__ __forceinline__ bool hull_vertex(V3 xw, uint32_t vid, const Obb& box, F4 qBox, V3 rel, float hullR,
                                            float boxR, const StepParams& P, HullVtx& out) {
  const V3 loc = q_rotate(q_conj(qBox), xw - box.c);
  const V3 cl = clamp_box(loc, box.e);
  V3 nL, bL;
  float depth;
  const bool inside = loc.x >= -box.e.x && loc.y >= -box.e.y && loc.z >= -box.e.z && loc.x <= box.e.x &&
                      loc.y <= box.e.y && loc.z <= box.e.z;
  if (inside) {
    // exit through the nearest face
    const V3 room = box.e - vabs(loc);
    const int j = min_axis(room);
    nL = v3s(0);
    setcomp(nL, j, comp(loc, j) >= 0.0f ? 1.0f : -1.0f);
    bL = loc;
    setcomp(bL, j, comp(nL, j) * comp(box.e, j));
    depth = -comp(room, j);
  } else {
    const V3 d = loc - cl;
    const float dist = length(d);
    if (dist < 1e-8f) return false;
    nL = d / dist;
    if (dist > detect_margin(q_rotate(qBox, nL), rel, P, hullR, boxR)) return false;
    bL = cl;
    depth = dist;
  }
  out.xHull = xw;
  out.xBox = q_rotate(qBox, bL) + box.c;
  out.n = q_rotate(qBox, nL);
  out.depth = depth;
  out.key = (4u << 24) | (vid & 0x00FFFFFFu);
  return true;
}

struct HullCtx {
  const F4* verts;   // this hull's vertices (collider-local)
  uint32_t count;
  F4 qH, qBox;
  V3 pH;
  Obb box;
  V3 rel;
  float hullR, boxR;
};
__device__ __forceinline__ bool hull_eval(const HullCtx& H, uint32_t v, const StepParams& P, HullVtx& out) {
  return hull_vertex(H.pH + q_rotate(H.qH, xyz(H.verts[v])), v, H.box, H.qBox, H.rel, H.hullR, H.boxR, P, out);
}

// The deepest vertex, then up to kMaxContacts - 1 more within 0.75 margin of
// its depth along the same face normal, each maximizing its least tangential
// distance to the chosen ones (ties within tolerance go to the lower vertex).
// Returns the number chosen (0: no contact); lane j < count ends holding
// chosen vertex j.  `nOut` = the deepest vertex's box -> hull normal.
__device__ __forceinline__ int hull_box(const HullCtx& H, const StepParams& P, F4* lds, HullVtx& mine, V3& nOut) {
  const int k = row_lane();
  bool have = false;
  float bDepth = 0.0f;
  uint32_t bKey = 0, bV = 0;
  for (uint32_t base = 0; base < H.count; base += kRow) {
    HullVtx c;
    c.depth = 0.0f;
    c.key = 0;
    const bool ok = base + k < H.count && hull_eval(H, base + k, P, c);
    const uint32_t okBits = row_bits(ok);
    for (int j = 0; j < kRow && base + j < H.count; j++) {
      const float dj = row_get(c.depth, j);
      const uint32_t kj = row_get(c.key, j);
      if (((okBits >> j) & 1u) &&
          (!have || dj < bDepth - 1e-8f || (fabsf(dj - bDepth) <= 1e-8f && kj < bKey))) {
        have = true;
        bDepth = dj;
        bKey = kj;
        bV = base + j;
      }
    }
  }
  if (!have) return 0;
  HullVtx best;
  hull_eval(H, bV, P, best);
  nOut = best.n;
  // chosen vertices: world position in xyz, vertex id in w
  if (k == 0) lds[0] = f4(best.xHull, as_float(bV));
  lds_wave_sync();
  int nc = 1;
  const float band = 0.75f * P.collisionMargin;
  while (nc < kMaxContacts) {
    bool haveNext = false;
    float spreadBest = -1.0f;
    uint32_t nKey = 0, nV = 0;
    for (uint32_t base = 0; base < H.count; base += kRow) {
      HullVtx c;
      c.key = 0;
      bool ok = base + k < H.count && hull_eval(H, base + k, P, c);
      if (ok && (c.depth > best.depth + band || dot(c.n, best.n) < 0.98f)) ok = false;
      float spread = kHuge;
      if (ok) {
        for (int j = 0; j < nc; j++) {
          const F4 ch = lds[j];
          if (base + k == as_uint(ch.w)) {
            ok = false;
            break;
          }
          V3 d = c.xHull - xyz(ch);
          d = d - best.n * dot(best.n, d);
          spread = fminf(spread, dot(d, d));
        }
      }
      const uint32_t okBits = row_bits(ok);
      for (int j = 0; j < kRow && base + j < H.count; j++) {
        const float sj = row_get(spread, j);
        const uint32_t kj = row_get(c.key, j);
        if (((okBits >> j) & 1u) &&
            (!haveNext || sj > spreadBest + 1e-10f || (fabsf(sj - spreadBest) <= 1e-10f && kj < nKey))) {
          haveNext = true;
          spreadBest = sj;
          nKey = kj;
          nV = base + j;
        }
      }
    }
    if (!haveNext || spreadBest < kMergeSq) break;
    HullVtx nx;
    hull_eval(H, nV, P, nx);
    lds_wave_sync();
    if (k == 0) lds[nc] = f4(nx.xHull, as_float(nV));
    lds_wave_sync();
    nc++;
  }
  if (k < nc) hull_eval(H, as_uint(lds[k].w), P, mine);
  lds_wave_sync();
  return nc;
}

// everything the row knows about its collider pair
struct PairView {
  uint32_t ca, cb, ba, bb, ta, tb;
  V3 pA, pB;   // owner positions
  F4 oA, oB;   // owner rotations
  V3 cA, cB;   // collider centers
  F4 qA, qB;   // collider rotations
  V3 vA, vB;   // velocities of the collider centers
  F4 shA, shB; // size xyz, signed bounding radius
};

__device__ __forceinline__ PairView load_pair(const DeviceState& S, uint32_t p) {
  PairView v;
  const U2 pr = S.pairs[p];
  v.ca = pr.x;
  v.cb = pr.y;
  v.ba = S.colliderOwner[v.ca];
  v.bb = S.colliderOwner[v.cb];
  v.ta = S.colliderShapeType[v.ca];
  v.tb = S.colliderShapeType[v.cb];
  v.pA = xyz(S.posLin[v.ba]);
  v.pB = xyz(S.posLin[v.bb]);
  v.oA = S.posAng[v.ba];
  v.oB = S.posAng[v.bb];
  v.cA = v.pA + q_rotate(v.oA, xyz(S.colliderLocalPos[v.ca]));
  v.cB = v.pB + q_rotate(v.oB, xyz(S.colliderLocalPos[v.cb]));
  v.qA = q_mul(v.oA, S.colliderLocalRot[v.ca]);
  v.qB = q_mul(v.oB, S.colliderLocalRot[v.cb]);
  v.vA = xyz(S.velLin[v.ba]) + cross(xyz(S.velAng[v.ba]), v.cA - v.pA);
  v.vB = xyz(S.velLin[v.bb]) + cross(xyz(S.velAng[v.bb]), v.cB - v.pB);
  v.shA = S.colliderShape[v.ca];
  v.shB = S.colliderShape[v.cb];
  return v;
}

// ---- round shapes: sphere / capsule pairs ----
// Fills the lane's hit (at most 3 for capsule-box, otherwise 1 on lane 0)
// and the normal from B toward A; returns false when nothing is in range.
__device__ __forceinline__ bool round_pair(const PairView& V, const Obb& A, const Obb& B, const StepParams& P, Hit& h, V3& nrm,
                           float& matchRadius) {
  const int k = row_lane();
  const uint32_t kindA = V.ta & kShapeKindMask, kindB = V.tb & kShapeKindMask;
  const float radA = V.shA.x * 0.5f, radB = V.shB.x * 0.5f;
  h.on = false;
  h.slot = 0;
  h.key = 0;
  if (kindA == 3u || kindB == 3u) {
    // capsule (segment along local z, length size.x, radius size.y) against
    // a sphere, a capsule or a box; points and normal in the capsule's frame
    // first (xA on the capsule), swapped to (A, B) at the end
    const bool capIsA = kindA == 3u;
    const F4 sCap = pick(capIsA, V.shA, V.shB), sOth = pick(capIsA, V.shB, V.shA);
    const F4 qCap = pick(capIsA, V.qA, V.qB), qOth = pick(capIsA, V.qB, V.qA);
    const V3 cCap = pick(capIsA, A.c, B.c), cOth = pick(capIsA, B.c, A.c);
    const float rCap = sCap.y;
    const V3 axis = q_rotate(qCap, v3(0, 0, 1));
    const V3 s0 = cCap - axis * (sCap.x * 0.5f), s1 = cCap + axis * (sCap.x * 0.5f);
    const uint32_t kindO = capIsA ? kindB : kindA;
    const V3 relO = pick(capIsA, V.vB - V.vA, V.vA - V.vB);
    V3 nCap = v3(0, 0, 1);  // capsule -> other
    V3 onCap = v3s(0), onOth = v3s(0);
    bool any = false;
    if (kindO == 1u || kindO == 3u) {
      // one closest-point contact (all lanes compute it; lane 0 keeps it)
      V3 qc, qo;
      float rO;
      if (kindO == 1u) {
        rO = sOth.x * 0.5f;
        qc = nearest_on_segment(cOth, s0, s1);
        qo = cOth;
      } else {
        rO = sOth.y;
        const V3 axO = q_rotate(qOth, v3(0, 0, 1));
        nearest_between_segments(s0, s1, cOth - axO * (sOth.x * 0.5f), cOth + axO * (sOth.x * 0.5f), qc, qo);
      }
      const V3 d = qo - qc;
      const float dist = length(d);
      const V3 nT = dist > 1e-6f ? d / dist : v3(0, 0, 1);
      if (dist <= rCap + rO + detect_margin(nT, relO, P, sCap.w, sOth.w)) {
        nCap = nT;
        onCap = qc + nCap * rCap;
        onOth = qo - nCap * rO;
        any = true;
        h.on = k == 0;
      }
    } else {
      // box: three seeds along the segment (lanes 0-2), each alternating
      // box clamp <-> segment projection five times
      const F4 qInv = q_conj(qOth);
      const V3 hb = xyz(sOth) * 0.5f;
      bool live = false;
      V3 pc = v3s(0), pb = v3s(0), ns = v3s(0);
      if (k < 3) {
        V3 q = s0 + (s1 - s0) * ((float)k * 0.5f);
        V3 bw = q;
        for (int it = 0; it < 5; it++) {
          bw = q_rotate(qOth, clamp_box(q_rotate(qInv, q - cOth), hb)) + cOth;
          q = nearest_on_segment(bw, s0, s1);
        }
        const V3 d = bw - q;
        const float dist = length(d);
        if (dist < 1e-6f) {
          // segment inside the box: out through the least-penetrated face
          const V3 lb = q_rotate(qInv, q - cOth);
          const V3 room = hb - vabs(lb);
          const int j = min_axis(room);
          V3 nl = v3s(0);
          setcomp(nl, j, comp(lb, j) >= 0.0f ? 1.0f : -1.0f);
          ns = -q_rotate(qOth, nl);
          bw = q - ns * comp(room, j);
          live = true;
        } else {
          ns = d / dist;
          live = !(dist - rCap > detect_margin(ns, relO, P, sCap.w, sOth.w));
        }
        pc = q + ns * rCap;
        pb = bw;
      }
      // seeds in order: a seed whose capsule point lies within 0.4 r + 5 cm
      // of a kept one is dropped; the normal is the sum of the kept ones'
      const float near = 0.4f * rCap + 0.05f;
      const V3 c0 = row_get(pc, 0), c1 = row_get(pc, 1), c2 = row_get(pc, 2);
      const uint32_t liveBits = row_bits(live);
      const bool keep0 = (liveBits & 1u) != 0u;
      const bool keep1 = (liveBits & 2u) != 0u && !(keep0 && length(c0 - c1) < near);
      const bool keep2 = (liveBits & 4u) != 0u && !(keep0 && length(c0 - c2) < near) &&
                         !(keep1 && length(c1 - c2) < near);
      V3 nSum = v3s(0);
      if (keep0) nSum += row_get(ns, 0);
      if (keep1) nSum += row_get(ns, 1);
      if (keep2) nSum += row_get(ns, 2);
      const int nk = (int)keep0 + (int)keep1 + (int)keep2;
      int mySlot = -1;
      if (k == 0 && keep0) mySlot = 0;
      if (k == 1 && keep1) mySlot = (int)keep0;
      if (k == 2 && keep2) mySlot = (int)keep0 + (int)keep1;
      if (nk > 0) {
        nCap = normalize(nSum);
        any = true;
      }
      onCap = pc;
      onOth = pb;
      h.on = mySlot >= 0;
      h.slot = mySlot >= 0 ? mySlot : 0;
      h.key = (uint32_t)k;
    }
    if (!any) return false;
    nrm = pick(capIsA, -nCap, nCap);
    h.xA = pick(capIsA, onCap, onOth);
    h.xB = pick(capIsA, onOth, onCap);
    matchRadius = 0.6f * (rCap + 0.2f);
    return true;
  }
  matchRadius = 0.5f * fmaxf(radA, radB);
  if (kindA == 1u && kindB == 1u) {
    const V3 d = A.c - B.c;
    const float dist = length(d);
    const V3 nT = dist > 1e-9f ? d / dist : v3(0, 0, 1);
    if (dist > radA + radB + detect_margin(nT, V.vA - V.vB, P, radA, radB)) return false;
    nrm = nT;
    h.xA = A.c - nrm * radA;
    h.xB = B.c + nrm * radB;
    h.on = k == 0;
    return true;
  }
  // sphere vs box
  const bool sphIsA = kindA == 1u;
  const float r = sphIsA ? radA : radB;
  const Obb box = pick(sphIsA, B, A);
  const F4 qBox = pick(sphIsA, V.qB, V.qA);
  const V3 sc = pick(sphIsA, A.c, B.c);
  const V3 loc = q_rotate(q_conj(qBox), sc - box.c);
  const V3 cl = clamp_box(loc, box.e);
  V3 nL, onL;
  if (cl.x == loc.x && cl.y == loc.y && cl.z == loc.z) {
    // center inside: out through the least-penetrated face
    const V3 room = box.e - vabs(loc);
    const int j = min_axis(room);
    nL = v3s(0);
    setcomp(nL, j, comp(loc, j) >= 0.0f ? 1.0f : -1.0f);
    onL = loc;
    setcomp(onL, j, comp(nL, j) * comp(box.e, j));
  } else {
    const V3 d = loc - cl;
    const float dist = length(d);
    const V3 nLT = d / fmaxf(dist, 1e-9f);
    const V3 vS = pick(sphIsA, V.vA, V.vB), vX = pick(sphIsA, V.vB, V.vA);
    if (dist > r + detect_margin(q_rotate(qBox, nLT), vS - vX, P, r, sphIsA ? V.shB.w : V.shA.w)) return false;
    nL = nLT;
    onL = cl;
  }
  const V3 nW = q_rotate(qBox, nL);
  const V3 onBox = q_rotate(qBox, onL) + box.c;
  const V3 onSph = sc - nW * r;
  nrm = pick(sphIsA, nW, -nW);
  h.xA = pick(sphIsA, onSph, onBox);
  h.xB = pick(sphIsA, onBox, onSph);
  h.on = k == 0;
  return true;
}

// ---- torus pairs (30_narrowphase.metal:292-336, 749-965) ----
// The torus is the set of points within r of its spine, the circle of radius
// R about the collider's local z axis through its center.
struct TorusGeo {
  V3 c, axis, u, v;
  float R, r;
};
__device__ __forceinline__ TorusGeo torus_at(V3 p, F4 q, float R, float r) {
  TorusGeo t;
  t.c = p;
  t.axis = q_rotate(q, v3(0.0f, 0.0f, 1.0f));
  const bool xMajor = fabsf(t.axis.x) > fabsf(t.axis.z);
  t.u = normalize(pick(xMajor, v3(-t.axis.y, t.axis.x, 0.0f), v3(0.0f, -t.axis.z, t.axis.y)));
  t.v = cross(t.axis, t.u);
  t.R = R;
  t.r = r;
  return t;
}
// nearest spine point (the spine point along u for points on the axis)
__device__ __forceinline__ V3 spine_nearest(const TorusGeo& t, V3 p) {
  V3 d = p - t.c;
  d = d - t.axis * dot(t.axis, d);
  const float l = length(d);
  return pick(l < 1e-6f, t.c + t.u * t.R, t.c + d / l * t.R);
}

// What the torus is tested against: a box (nearest point by clamping), a
// capsule (its segment) or another torus (its spine).
struct TorusOther {
  bool isBox, isCap;
  V3 p, half, s0, s1;
  F4 q, qc;
  TorusGeo t2;
  float r;  // tube / capsule radius, 0 for a box
};
__device__ __forceinline__ V3 torus_other_nearest(const TorusOther& O, V3 q) {
  if (O.isBox) return q_rotate(O.q, clamp_box(q_rotate(O.qc, q - O.p), O.half)) + O.p;
  if (O.isCap) return nearest_on_segment(q, O.s0, O.s1);
  return spine_nearest(O.t2, q);
}

// Lane k evaluates spine samples k and k + 16 (angles 2 pi i / 32 from the
// cos / sin table); the four greedy picks are row arg-mins of the sample gap
// (lowest sample on ties), each masking the samples within 3 of it; lane j
// then refines pick j by two alternating projections and a ballot packs the
// kept hits in pick order.  The sphere case is closed form on lane 0.
// Hits carry (xA, xB) on colliders A / B; `nrm` points from B toward A.
__device__ __forceinline__ bool torus_pair(const PairView& V, const StepParams& P, const float* trig, Hit& h, V3& nrm,
                                           float& matchRadius) {
  const int k = row_lane();
  const uint32_t kindA = V.ta & kShapeKindMask, kindB = V.tb & kShapeKindMask;
  const bool tIsA = kindA == 2u;
  const F4 shT = pick(tIsA, V.shA, V.shB), shO = pick(tIsA, V.shB, V.shA);
  const V3 pO = pick(tIsA, V.cB, V.cA);
  const TorusGeo T = torus_at(pick(tIsA, V.cA, V.cB), pick(tIsA, V.qA, V.qB), shT.x, shT.y);
  const uint32_t other = tIsA ? kindB : kindA;
  const float detectM = fmaxf(P.collisionMargin, 0.8f * T.r);  // fixed reach: a fraction of the tube radius
  matchRadius = 1.1f * shT.x;
  bool kept = false;
  V3 xT = v3s(0), xO = v3s(0), n = v3s(0);
  float sep = 0.0f;
  uint32_t feat = 0;
  if (other == 1u) {
    if (k == 0) {
      const float rs = shO.x * 0.5f;
      const V3 spine = spine_nearest(T, pO);
      const V3 d = pO - spine;
      const float dist = length(d);
      if (dist <= T.r + rs + detectM) {
        const V3 nts = pick(dist > 1e-6f, d / dist, v3(0.0f, 0.0f, 1.0f));  // torus -> sphere
        xT = spine + nts * T.r;
        xO = pO - nts * rs;
        n = -nts;
        sep = dist - T.r - rs;
        kept = true;
      }
    }
  } else {
    TorusOther O;
    O.isBox = other == 0u;
    O.isCap = other == 3u;
    O.p = pO;
    O.q = pick(tIsA, V.qB, V.qA);
    O.qc = q_conj(O.q);
    O.half = xyz(shO) * 0.5f;
    O.r = 0.0f;
    O.s0 = O.s1 = v3s(0);
    O.t2 = TorusGeo{};
    if (O.isCap) {
      const float ch = shO.x * 0.5f;
      const V3 ax = q_rotate(O.q, v3(0.0f, 0.0f, 1.0f));
      O.s0 = pO - ax * ch;
      O.s1 = pO + ax * ch;
      O.r = shO.y;
    } else if (!O.isBox) {
      O.t2 = torus_at(pO, O.q, shO.x, shO.y);
      O.r = O.t2.r;
    }
    const V3 q0 = T.c + (T.u * trig[2 * k] + T.v * trig[2 * k + 1]) * T.R;
    const V3 q1 = T.c + (T.u * trig[2 * k + 32] + T.v * trig[2 * k + 33]) * T.R;
    const float d0 = length(q0 - torus_other_nearest(O, q0));
    const float d1 = length(q1 - torus_other_nearest(O, q1));
    bool u0 = false, u1 = false;
    int mine = -1, picks = 0;
    for (int pk = 0; pk < 4; pk++) {
      float bd = kHuge;
      int bi = -1;
      if (!u0 && d0 < bd) { bd = d0; bi = k; }
      if (!u1 && d1 < bd) { bd = d1; bi = k + 16; }
      for (int o = 8; o > 0; o >>= 1) {
        const float d2 = __shfl_xor(bd, o, kRow);
        const int i2 = __shfl_xor(bi, o, kRow);
        if (i2 >= 0 && (bi < 0 || d2 < bd || (d2 == bd && i2 < bi))) {
          bd = d2;
          bi = i2;
        }
      }
      bd = row_get(bd, 0);
      bi = row_get(bi, 0);
      if (bi < 0 || bd - T.r - O.r > detectM) break;
      int c0 = abs(k - bi), c1 = abs(k + 16 - bi);
      c0 = min(c0, 32 - c0);
      c1 = min(c1, 32 - c1);
      u0 = u0 || c0 <= 3;
      u1 = u1 || c1 <= 3;
      if (k == pk) mine = bi;
      picks++;
    }
    if (k < picks) {
      V3 q = T.c + (T.u * trig[2 * mine] + T.v * trig[2 * mine + 1]) * T.R;
      V3 b = torus_other_nearest(O, q);
      for (int r2 = 0; r2 < 2; r2++) {
        q = spine_nearest(T, b);
        b = torus_other_nearest(O, q);
      }
      const V3 d = q - b;
      const float dist = length(d);
      V3 xo = b;
      if (dist < 1e-6f) {
        if (O.isBox) {
          // spine point inside the box: out through the nearest face
          const V3 lb = q_rotate(O.qc, q - O.p);
          const V3 pen = O.half - v3(fabsf(lb.x), fabsf(lb.y), fabsf(lb.z));
          const int ax = min_axis(pen);
          V3 nl = v3s(0);
          setcomp(nl, ax, comp(lb, ax) >= 0.0f ? 1.0f : -1.0f);
          n = q_rotate(O.q, nl);
          xo = q + n * comp(pen, ax);
        } else {
          const V3 other2 = pick(O.isCap, normalize(O.s1 - O.s0), O.t2.axis);
          n = normalize(cross(T.axis, other2) + v3(1e-4f, 0.0f, 0.0f));
        }
        kept = true;
      } else {
        kept = !(dist - T.r - O.r > detectM);
        n = d / dist;  // other -> torus
      }
      xT = q - n * T.r;
      xO = pick(O.isBox, xo, xo + n * O.r);
      sep = dist - T.r - O.r;
      feat = (uint32_t)mine;
    }
  }
  const uint32_t bits = row_bits(kept);
  if (bits == 0u) return false;
  // deepest hit (first in pick order on ties) sets the normal
  bool ok = kept && sep < kHuge;
  float best = ok ? sep : kHuge;
  int lane = ok ? k : 16;
  for (int o = 8; o > 0; o >>= 1) {
    const bool ok2 = __shfl_xor((int)ok, o, kRow) != 0;
    const float s2 = __shfl_xor(best, o, kRow);
    const int l2 = __shfl_xor(lane, o, kRow);
    if (ok2 && (!ok || s2 < best || (s2 == best && l2 < lane))) {
      ok = true;
      best = s2;
      lane = l2;
    }
  }
  lane = row_get(lane, 0);
  if (lane >= 16) lane = __ffs(bits) - 1;
  const V3 nDeep = row_get(n, lane);
  nrm = pick(tIsA, nDeep, -nDeep);
  h.on = kept;
  h.slot = __popc(bits & ((1u << k) - 1u));
  h.xA = pick(tIsA, xT, xO);
  h.xB = pick(tIsA, xO, xT);
  h.key = feat;
  return true;
}

// Kinematic spinner surface motion as a tangential target
// (30_narrowphase.metal:274-284): each spinning body's surface displacement
// over the step at the contact points (offsets from the body origins),
// pre-divided by 1 - alpha because the solver discounts C0 by alpha.
__device__ __forceinline__ void spin_shift(const DeviceState& S, const StepParams& P, uint32_t ba, uint32_t bb, V3 offA,
                                           V3 offB, V3 t1, V3 t2, F4& C0) {
  const F4 wA = S.spinVel[ba], wB = S.spinVel[bb];
  const V3 sA = pick(wA.w != 0.0f, cross(xyz(wA), offA) * P.dt, v3s(0));
  const V3 sB = pick(wB.w != 0.0f, cross(xyz(wB), offB) * P.dt, v3s(0));
  const V3 rel = sB - sA;
  const float kk = 1.0f / fmaxf(1.0f - P.alpha, 0.01f);
  C0.y = C0.y - dot(t1, rel) * kk;
  C0.z = C0.z - dot(t2, rel) * kk;
}

__global__ void __launch_bounds__(256) k_np_collide(DeviceState S, StepParams P) {
  __shared__ F4 scratch[kRowsPerGroup][kPolyMax];
  F4* lds = scratch[threadIdx.x >> 4];
  const int k = row_lane();
  const uint32_t numPairs = S.counters[CTR_PAIRS];
  const uint32_t rowsTotal = gridDim.x * kRowsPerGroup;
  for (uint32_t p = blockIdx.x * kRowsPerGroup + (threadIdx.x >> 4); p < numPairs; p += rowsTotal) {
    const PairView V = load_pair(S, p);
    ManifoldRec& M = S.manifolds[p];
    const Obb A = make_obb(V.cA, V.qA, xyz(V.shA));
    const Obb B = make_obb(V.cB, V.qB, xyz(V.shB));
    const uint32_t kindA = V.ta & kShapeKindMask, kindB = V.tb & kShapeKindMask;
    const bool hullA = kindA == 4u, hullB = kindB == 4u;
    const bool torusPair = kindA == 2u || kindB == 2u;
    const bool roundPair = !torusPair && (kindA == 1u || kindB == 1u || kindA == 3u || kindB == 3u);
    if (k == 0) M.colliderPair = U4{V.ca, V.cb, 0, 0};

    Hit h;
    h.on = false;
    h.slot = 0;
    h.key = 0;
    V3 nrm = v3(0, 0, 1);
    bool touching = false;
    float matchRadius = 0.0f;  // round pairs: proximity bound of the warm-start match
    // convex hulls collide with boxes only (30_narrowphase.metal:404-410)
    const bool refused = (hullA || hullB) && ((hullA && hullB) || (hullA ? kindB : kindA) != 0u);
    if (refused) {
      touching = false;
    } else if (torusPair) {
      touching = torus_pair(V, P, S.torusTrig, h, nrm, matchRadius);
    } else if (roundPair) {
      touching = round_pair(V, A, B, P, h, nrm, matchRadius);
    } else if (hullA || hullB) {
      HullCtx H;
      const uint32_t ch = hullA ? V.ca : V.cb;
      const U2 range = S.colliderHullRange[ch];
      H.verts = S.hullVerts + range.x;
      H.count = range.y;
      H.qH = pick(hullA, V.qA, V.qB);
      H.qBox = pick(hullA, V.qB, V.qA);
      H.pH = pick(hullA, A.c, B.c);
      H.box = pick(hullA, B, A);
      H.rel = pick(hullA, V.vA - V.vB, V.vB - V.vA);
      H.hullR = fabsf(hullA ? V.shA.w : V.shB.w);
      H.boxR = fabsf(hullA ? V.shB.w : V.shA.w);
      HullVtx mine;
      V3 nBoxToHull;
      const int nc = hull_box(H, P, lds, mine, nBoxToHull);
      touching = nc > 0;
      if (touching) {
        h.on = k < nc;
        h.slot = k;
        h.xA = pick(hullA, mine.xHull, mine.xBox);
        h.xB = pick(hullA, mine.xBox, mine.xHull);
        h.key = mine.key;
        nrm = pick(hullA, nBoxToHull, -nBoxToHull);
      }
    } else {
      touching = box_box(A, B, V.vB - V.vA, spec_cap(V.shA.w, V.shB.w, P.collisionMargin), P, lds, h, nrm);
    }
    const uint32_t count = __popc(row_bits(touching && h.on));
    if (count == 0u) {
      if (k == 0) M.header = U4{V.ba, V.bb, 0, 0};
      continue;
    }

    // ---- manifold record and warm start ----
    V3 t1, t2;
    orthonormal(nrm, t1, t2);
    const F4 fA = S.colliderFriction[V.ca], fB = S.colliderFriction[V.cb];
    const bool worldA = (roundPair || torusPair) && (V.ta & kShapeWorldRound) != 0u;
    const bool worldB = (roundPair || torusPair) && (V.tb & kShapeWorldRound) != 0u;
    if (k == 0) {
      const uint32_t flags = 1u | (worldA ? 2u : 0u) | (worldB ? 4u : 0u);
      M.header = U4{V.ba, V.bb, count, flags};
      M.basisN = f4(nrm, combine_friction(fA.y, fB.y, P.frictionCombineMode));
      M.basisT1 = f4(t1, combine_friction(fA.x, fB.x, P.frictionCombineMode));
    }
    int prev = k == 0 ? pair_map_find(S, P.mapCapacity, V.ca, V.cb) : 0;
    prev = row_get(prev, 0);
    if (!h.on) continue;
    // anchors: body-local, or world offsets for world-round colliders
    V3 ancA = pick(worldA, h.xA - V.pA, q_rotate(q_conj(V.oA), h.xA - V.pA));
    V3 ancB = pick(worldB, h.xB - V.pB, q_rotate(q_conj(V.oB), h.xB - V.pB));
    V3 lam = v3s(0), pen = v3s(0);
    float stick = 0.0f;
    if (prev >= 0) {
      const ManifoldRec& Q = S.prevManifolds[prev];
      const uint32_t qn = Q.header.z;
      if (roundPair || torusPair) {
        // nearest previous anchor on A within the match radius; its
        // tangential multiplier is carried into this frame's basis
        float bestD = matchRadius;
        int bj = -1;
        for (uint32_t j = 0; j < qn; j++) {
          const float dj = length(xyz(Q.c[j].rA) - ancA);
          if (dj < bestD) {
            bestD = dj;
            bj = (int)j;
          }
        }
        if (bj >= 0) {
          lam = xyz(Q.c[bj].lambda);
          pen = xyz(Q.c[bj].penalty);
          const V3 u1 = xyz(Q.basisT1), u2 = cross(xyz(Q.basisN), u1);
          const V3 ft = u1 * lam.y + u2 * lam.z;
          lam.y = dot(t1, ft);
          lam.z = dot(t2, ft);
          // a sticking anchor is reused unless it is a world offset of a
          // rigid round shape (capsule pairs: unless either is a world offset)
          // (torus contacts never stick: their anchors are this frame's hits)
          const bool capsule = kindA == 3u || kindB == 3u;
          const bool reuse = !torusPair && (capsule ? (!worldA && !worldB)
                                                    : ((V.shA.w < 0.0f || !worldA) && (V.shB.w < 0.0f || !worldB)));
          if (reuse) {
            stick = Q.c[bj].rB.w;
            if (stick != 0.0f) {
              ancA = xyz(Q.c[bj].rA);
              ancB = xyz(Q.c[bj].rB);
            }
          }
        }
      } else {
        for (uint32_t j = 0; j < qn; j++) {
          if (as_uint(Q.c[j].rA.w) != h.key) continue;
          lam = xyz(Q.c[j].lambda);
          pen = xyz(Q.c[j].penalty);
          stick = Q.c[j].rB.w;
          if (stick != 0.0f) {
            ancA = xyz(Q.c[j].rA);
            ancB = xyz(Q.c[j].rB);
          }
          break;
        }
      }
    }
    // torus contacts measure the gap between this frame's hit points directly
    const V3 wA = pick(torusPair, h.xA, pick(worldA, V.pA + ancA, xform(V.pA, V.oA, ancA)));
    const V3 wB = pick(torusPair, h.xB, pick(worldB, V.pB + ancB, xform(V.pB, V.oB, ancB)));
    const V3 gap = wA - wB;
    const V3 floorK = contact_penalty_floor(S, V.ba, V.bb, P);
    ContactRec& c = M.c[h.slot];
    c.rA = f4(ancA, as_float(h.key));
    c.rB = f4(ancB, stick);
    c.C0 = f4(v3(dot(nrm, gap) + P.collisionMargin, dot(t1, gap), dot(t2, gap)), 0.0f);
    if (S.spinVel) spin_shift(S, P, V.ba, V.bb, wA - V.pA, wB - V.pB, t1, t2, c.C0);
    c.lambda = f4(lam * (P.alpha * P.gamma), 0.0f);
    c.penalty = f4(v3(clampf(pen.x * P.gamma, floorK.x, kPenaltyMax), clampf(pen.y * P.gamma, floorK.y, kPenaltyMaxT),
                      clampf(pen.z * P.gamma, floorK.z, kPenaltyMaxT)),
                   0.0f);
  }
}

// Kinematic spinners (GPUSolver.swift:4418-4427): before the frame each
// spinning body's orientation becomes normalize(dq q) for each of its
// spinners in order, dq the host-computed rotation by omega dt about the
// spinner axis.  One thread per spinning body.
__global__ void __launch_bounds__(64) k_spinners(DeviceState S, const uint32_t* bodies, const uint32_t* start,
                                                 const F4* dq, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t b = bodies[t];
  F4 q = S.posAng[b];
  for (uint32_t k = start[t]; k < start[t + 1]; k++) {
    const F4 r = q_mul(dq[k], q);
    const float sc = 1.0f / sqrtf(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
    q = F4{r.x * sc, r.y * sc, r.z * sc, r.w * sc};
  }
  S.posAng[b] = q;
}

}  // namespace

void launch_spinners(hipStream_t st, const DeviceState& S, const uint32_t* bodies, const uint32_t* start,
                     const F4* dq, uint32_t n) {
  hipLaunchKernelGGL(k_spinners, dim3((n + 63) / 64), dim3(64), 0, st, S, bodies, start, dq, n);
}

void launch_narrowphase(hipStream_t st, const DeviceState& S, const StepParams& P, uint32_t gridCap) {
  uint32_t blocks = (P.maxPairs + kRowsPerGroup - 1) / kRowsPerGroup;
  if (blocks > gridCap) blocks = gridCap;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_np_collide, dim3(blocks), dim3(kRowsPerGroup * kRow), 0, st, S, P);
}

}  // namespace avbd
