// common.hpp -- device math, record layouts and step parameters shared by the
// gfx950 AVBD kernels and the host runtime that fills them.
//
// Records are 16-byte granular so a gather of one record field is a single
// global_load_dwordx4.  Field meanings follow the reference's wire records
// (Sources/PhysicsAVBD/GPU/GPUTypes.swift:80-173) because the solver math
// indexes them the same way; the layouts are this backend's own.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AVBD_HD __host__ __device__ __forceinline__

namespace avbd {

// ---------------------------------------------------------------------------
// constants (00_common.metal:11-45)
// ---------------------------------------------------------------------------
constexpr float kPenaltyMin = 1.0f;
constexpr float kPenaltyMax = 1.0e10f;
constexpr float kPenaltyMaxT = 1.0e6f;   // tangential / soft normal cap
constexpr float kStickThresh = 1.0e-5f;
constexpr int kMaxColors = 64;
constexpr int kMaxContacts = 8;
constexpr uint32_t kWorld = 0xFFFFFFFFu;

// force kinds packed in the top 4 bits of an adjacency entry
constexpr uint32_t FK_JOINT = 0, FK_SPRING = 1, FK_MANIFOLD = 2, FK_TET = 3,
                   FK_SOFT = 4, FK_MEMBRANE = 5, FK_BEND = 6,
                   FK_PSPRING = 7;  // particle-particle spring with a stamp-major slot
constexpr uint32_t kAdjShift = 28;
constexpr uint32_t kAdjIndexMask = 0x0FFFFFFFu;
// soft-contact entries: index in bits 0-23, bits 24-27 flag the participant
// slots that share the body's static color (read from the snapshot)
constexpr uint32_t kSoftIndexMask = 0x00FFFFFFu;
constexpr uint32_t kSoftSameShift = 24;
constexpr uint32_t kMaxSoftRecords = 1u << 24;
// partner id in a cloth slot: bit 31 = partner shares the body's static color
constexpr uint32_t kSnapBit = 0x80000000u;

// collider shape flags: low nibble kind, 0x10 particle, 0x20 world-round anchor
constexpr uint32_t kShapeKindMask = 0xFu;
constexpr uint32_t kShapeParticle = 0x10u;
constexpr uint32_t kShapeWorldRound = 0x20u;

// joint motor mode, bits 4-5 of the joint flags
constexpr uint32_t kMotorModeMask = 3u << 4;
constexpr uint32_t kMotorImplicitPD = 1u << 4;
constexpr uint32_t kMotorExplicitPD = 2u << 4;
constexpr uint32_t kMotorVelocity = 3u << 4;

// counters (single u32 buffer)
enum Counter : uint32_t {
  CTR_PAIRS = 0,
  CTR_SOFT = 1,
  CTR_GBAR = 2,            // grid barrier arrival counter (persistent sweep)
  CTR_GBAR_FAIL = 3,       // grid barrier spin bound exceeded
  CTR_PAIR_CANDIDATES = 4,
  CTR_SOFT_CANDIDATES = 5,
  CTR_COLOR_CONFLICTS = 6,
  CTR_ACTIVE_MANIFOLDS = 7,
  CTR_COLOR_CHANGED = 8,   // static/dynamic coloring fixed-point flag
  CTR_DYN_ADJ_TOTAL = 9,
  CTR_REPAIR_BAR = 10,     // grid barrier of the cooperative color repair
  CTR_REPAIR_PASSES = 11,  // passes the repair needed (diagnostic)
  CTR_DAT_PAIRS = 12,      // surface pair stream demand of this frame (V-T + E-E)
  CTR_DAT_OVERFLOW = 13,   // elements whose cover exceeded 4 cells on an axis
  CTR_DAT_NONFINITE = 14,  // non-finite element / query anchors
  CTR_SOFT_RT = 15,        // rigid-feature vs triangle records (prefix of the soft array)
  CTR_COLOR_BASE = 16,     // kMaxColors entries
  CTR_DAT_VT = 16 + 2 * kMaxColors,
  CTR_DAT_EE,
  CTR_WS_ACCEPT,           // learned warm start kept this frame (1) or fell back (0)
  CTR_WS_RES_PLAIN,        // residual at the plain start (f32 bits)
  CTR_WS_RES_LEARNED,      // residual at the learned start (f32 bits)
  CTR_LARGE_CELLS,         // hash buckets of more than 32 colliders this frame
  CTR_DAT_INVALID,         // invalid truncation anchors (non-finite or unresolvable planes)
  CTR_DAT_TRUNC,           // surface particles truncated by a Planar-DAT pass
  CTR_DAT_VT_DEGEN,        // coincident V-T anchors (fallback normal)
  CTR_DAT_EE_DEGEN,        // coincident / parallel E-E anchors (fallback normal)
  // residual-terminated frames: chunks of sweeps each end with a residual
  // check; once CTR_CONV_DONE is set every later sweep kernel of the frame
  // returns at entry
  CTR_CONV_DONE,           // 0 iterating, 1 residual <= tolerance, 2 iteration cap reached
  CTR_CONV_SEQ,            // residual checks completed this frame
  CTR_CONV_ITERS,          // sweeps run this frame
  CTR_CONV_RES,            // residual of the last check (f32 bits)
  // isotropic OGC limiter
  CTR_SOFT_ISO,            // tracked V-T / E-E records (after the rigid-feature ones)
  CTR_OGC,                 // truncations since the last refresh check
  CTR_OGC_ARRIVE,          // workgroups of the running refresh check that have read CTR_OGC
  CTR_OGC_TRUNC,           // truncations this frame
  CTR_OGC_REFRESH,         // refreshes fired this frame
  CTR_TOTAL = 16 + 2 * kMaxColors + 24
};

// Progress of a residual-terminated frame, written by the device into mapped
// host memory after every residual check (seq last, with system-scope release).
struct ConvHost {
  uint32_t seq, done, iters;
  float residual;
};

// ---------------------------------------------------------------------------
// small vector types
// ---------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
struct alignas(16) F4 {
  float x, y, z, w;
};
struct alignas(16) U4 {
  uint32_t x, y, z, w;
};

AVBD_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
AVBD_HD V3 v3s(float s) { return V3{s, s, s}; }
AVBD_HD V3 xyz(F4 a) { return V3{a.x, a.y, a.z}; }
AVBD_HD F4 f4(V3 a, float w) { return F4{a.x, a.y, a.z, w}; }
AVBD_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AVBD_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AVBD_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
AVBD_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
AVBD_HD V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
AVBD_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AVBD_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
AVBD_HD V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
AVBD_HD V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
AVBD_HD V3& operator*=(V3& a, float s) { a = a * s; return a; }
AVBD_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Error-free sum (Knuth's TwoSum): s = f32(a + b) and a + b = s + e exactly.
// Additions only, so contraction cannot fuse any of it.
AVBD_HD void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
AVBD_HD void two_sum3(V3 a, V3 b, V3& s, V3& e) {
  two_sum(a.x, b.x, s.x, e.x);
  two_sum(a.y, b.y, s.y, e.y);
  two_sum(a.z, b.z, s.z, e.z);
}
AVBD_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
AVBD_HD float length(V3 a) { return sqrtf(dot(a, a)); }
AVBD_HD float rsq(float x) { return 1.0f / sqrtf(x); }
AVBD_HD V3 normalize(V3 a) { return a * rsq(dot(a, a)); }
AVBD_HD V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
AVBD_HD V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
AVBD_HD V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
AVBD_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
AVBD_HD void setcomp(V3& a, int i, float v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}
AVBD_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// exponent-bits finiteness test (stays correct under any fp flags)
AVBD_HD bool finite_bits(float x) {
  union { float f; uint32_t u; } c; c.f = x;
  return (c.u & 0x7F800000u) != 0x7F800000u;
}
AVBD_HD bool finite3(V3 v) { return finite_bits(v.x) && finite_bits(v.y) && finite_bits(v.z); }
AVBD_HD float as_float(uint32_t u) { union { uint32_t u; float f; } c; c.u = u; return c.f; }
AVBD_HD uint32_t as_uint(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }

// ---------------------------------------------------------------------------
// row-major 3x3 (00_common.metal:50-83)
// ---------------------------------------------------------------------------
struct M3 {
  V3 r0, r1, r2;
};
AVBD_HD M3 m3_zero() { return M3{v3s(0), v3s(0), v3s(0)}; }
AVBD_HD M3 m3_diag(V3 d) { return M3{v3(d.x, 0, 0), v3(0, d.y, 0), v3(0, 0, d.z)}; }
AVBD_HD M3 m3_identity() { return m3_diag(v3s(1)); }
AVBD_HD V3 m3_mul(const M3& m, V3 v) { return V3{dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)}; }
AVBD_HD M3 m3_transpose(const M3& m) {
  return M3{v3(m.r0.x, m.r1.x, m.r2.x), v3(m.r0.y, m.r1.y, m.r2.y), v3(m.r0.z, m.r1.z, m.r2.z)};
}
AVBD_HD M3 m3_mulm(const M3& a, const M3& b) {
  M3 bt = m3_transpose(b);
  return M3{v3(dot(a.r0, bt.r0), dot(a.r0, bt.r1), dot(a.r0, bt.r2)),
            v3(dot(a.r1, bt.r0), dot(a.r1, bt.r1), dot(a.r1, bt.r2)),
            v3(dot(a.r2, bt.r0), dot(a.r2, bt.r1), dot(a.r2, bt.r2))};
}
AVBD_HD M3 m3_add(const M3& a, const M3& b) { return M3{a.r0 + b.r0, a.r1 + b.r1, a.r2 + b.r2}; }
AVBD_HD M3 m3_scale(const M3& a, float s) { return M3{a.r0 * s, a.r1 * s, a.r2 * s}; }
AVBD_HD M3 m3_outer(V3 a, V3 b) { return M3{b * a.x, b * a.y, b * a.z}; }
AVBD_HD M3 m3_skew(V3 r) { return M3{v3(0, -r.z, r.y), v3(r.z, 0, -r.x), v3(-r.y, r.x, 0)}; }
AVBD_HD V3 m3_col(const M3& m, int i) { return v3(comp(m.r0, i), comp(m.r1, i), comp(m.r2, i)); }
AVBD_HD M3 m3_diagonalize(const M3& m) {
  return m3_diag(v3(length(m3_col(m, 0)), length(m3_col(m, 1)), length(m3_col(m, 2))));
}

// ---------------------------------------------------------------------------
// quaternions (x,y,z imaginary, w real) -- 00_common.metal:88-124
// ---------------------------------------------------------------------------
AVBD_HD F4 q_mul(F4 a, F4 b) {
  return F4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
AVBD_HD F4 q_conj(F4 q) { return F4{-q.x, -q.y, -q.z, q.w}; }
AVBD_HD F4 q_inv(F4 q) {
  float n = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  return F4{-q.x / n, -q.y / n, -q.z / n, q.w / n};
}
AVBD_HD V3 q_rotate(F4 q, V3 v) {
  V3 qv = xyz(q);
  V3 t = cross(qv, v) * 2.0f;
  return v + t * q.w + cross(qv, t);
}
AVBD_HD M3 world_inertia(F4 q, V3 principal) {
  V3 c0 = q_rotate(q, v3(1, 0, 0)), c1 = q_rotate(q, v3(0, 1, 0)), c2 = q_rotate(q, v3(0, 0, 1));
  return m3_add(m3_add(m3_scale(m3_outer(c0, c0), principal.x), m3_scale(m3_outer(c1, c1), principal.y)),
                m3_scale(m3_outer(c2, c2), principal.z));
}
// a - b in tangent space, canonicalized to the w >= 0 hemisphere
AVBD_HD V3 q_sub(F4 a, F4 b) {
  F4 r = q_mul(a, q_inv(b));
  if (r.w < 0.0f) { r.x = -r.x; r.y = -r.y; r.z = -r.z; r.w = -r.w; }
  return v3(r.x, r.y, r.z) * 2.0f;
}
AVBD_HD F4 q_addw(F4 q, V3 w) {
  F4 dq = q_mul(F4{w.x, w.y, w.z, 0.0f}, q);
  F4 r{q.x + 0.5f * dq.x, q.y + 0.5f * dq.y, q.z + 0.5f * dq.z, q.w + 0.5f * dq.w};
  float s = rsq(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
  return F4{r.x * s, r.y * s, r.z * s, r.w * s};
}
AVBD_HD V3 xform(V3 p, F4 q, V3 v) { return q_rotate(q, v) + p; }

// orthonormal basis rows (n, t1, t2) -- 00_common.metal:224-228
AVBD_HD void orthonormal(V3 n, V3& t1, V3& t2) {
  t1 = fabsf(n.x) > fabsf(n.z) ? v3(-n.y, n.x, 0) : v3(0, -n.z, n.y);
  t1 = normalize(t1);
  t2 = cross(n, t1);
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned 6x6 LDL^T with pivot clamp (00_common.metal:137-221).
// Reads the lower triangle + diagonal of aLin / aAng and all of aCross
// (rows angular, columns linear).
// ---------------------------------------------------------------------------
AVBD_HD void solve6x6(const M3& aLin, const M3& aAng, const M3& aCross, V3 bLin, V3 bAng, V3& xLin,
                      V3& xAng) {
  const float s1 = rsq(fmaxf(aLin.r0.x, 1e-30f)), s2 = rsq(fmaxf(aLin.r1.y, 1e-30f)),
              s3 = rsq(fmaxf(aLin.r2.z, 1e-30f)), s4 = rsq(fmaxf(aAng.r0.x, 1e-30f)),
              s5 = rsq(fmaxf(aAng.r1.y, 1e-30f)), s6 = rsq(fmaxf(aAng.r2.z, 1e-30f));
  const float A21 = aLin.r1.x * s2 * s1;
  const float A31 = aLin.r2.x * s3 * s1, A32 = aLin.r2.y * s3 * s2;
  const float A41 = aCross.r0.x * s4 * s1, A42 = aCross.r0.y * s4 * s2, A43 = aCross.r0.z * s4 * s3;
  const float A51 = aCross.r1.x * s5 * s1, A52 = aCross.r1.y * s5 * s2, A53 = aCross.r1.z * s5 * s3,
              A54 = aAng.r1.x * s5 * s4;
  const float A61 = aCross.r2.x * s6 * s1, A62 = aCross.r2.y * s6 * s2, A63 = aCross.r2.z * s6 * s3,
              A64 = aAng.r2.x * s6 * s4, A65 = aAng.r2.y * s6 * s5;
  const float b1 = bLin.x * s1, b2 = bLin.y * s2, b3 = bLin.z * s3;
  const float b4 = bAng.x * s4, b5 = bAng.y * s5, b6 = bAng.z * s6;
  const float eps = 1e-6f;
  const float D1 = 1.0f;
  const float L21 = A21, L31 = A31, L41 = A41, L51 = A51, L61 = A61;
  const float D2 = fmaxf(1.0f - L21 * L21 * D1, eps);
  const float L32 = (A32 - L21 * L31 * D1) / D2;
  const float L42 = (A42 - L21 * L41 * D1) / D2;
  const float L52 = (A52 - L21 * L51 * D1) / D2;
  const float L62 = (A62 - L21 * L61 * D1) / D2;
  const float D3 = fmaxf(1.0f - (L31 * L31 * D1 + L32 * L32 * D2), eps);
  const float L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3;
  const float L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3;
  const float L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
  const float D4 = fmaxf(1.0f - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3), eps);
  const float L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4;
  const float L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
  const float D5 = fmaxf(1.0f - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4), eps);
  const float L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
  const float D6 =
      fmaxf(1.0f - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5), eps);
  const float y1 = b1;
  const float y2 = b2 - L21 * y1;
  const float y3 = b3 - L31 * y1 - L32 * y2;
  const float y4 = b4 - L41 * y1 - L42 * y2 - L43 * y3;
  const float y5 = b5 - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
  const float y6 = b6 - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;
  const float z1 = y1 / D1, z2 = y2 / D2, z3 = y3 / D3, z4 = y4 / D4, z5 = y5 / D5, z6 = y6 / D6;
  const float a3 = z6;
  const float a2 = z5 - L65 * a3;
  const float a1 = z4 - L54 * a2 - L64 * a3;
  const float l3 = z3 - L43 * a1 - L53 * a2 - L63 * a3;
  const float l2 = z2 - L32 * l3 - L42 * a1 - L52 * a2 - L62 * a3;
  const float l1 = z1 - L21 * l2 - L31 * l3 - L41 * a1 - L51 * a2 - L61 * a3;
  xLin = v3(l1 * s1, l2 * s2, l3 * s3);
  xAng = v3(a1 * s4, a2 * s5, a3 * s6);
}

// The same factorization restricted to the linear block: with aAng = I and
// aCross = 0 the angular rows decouple exactly, so particles solve 3x3.
AVBD_HD V3 solve3x3(float a00, float a10, float a11, float a20, float a21, float a22, V3 b) {
  const float s1 = rsq(fmaxf(a00, 1e-30f)), s2 = rsq(fmaxf(a11, 1e-30f)), s3 = rsq(fmaxf(a22, 1e-30f));
  const float L21 = a10 * s2 * s1, A31 = a20 * s3 * s1, A32 = a21 * s3 * s2;
  const float eps = 1e-6f;
  const float D2 = fmaxf(1.0f - L21 * L21, eps);
  const float L31 = A31;
  const float L32 = (A32 - L21 * L31) / D2;
  const float D3 = fmaxf(1.0f - (L31 * L31 + L32 * L32 * D2), eps);
  const float y1 = b.x * s1;
  const float y2 = b.y * s2 - L21 * y1;
  const float y3 = b.z * s3 - L31 * y1 - L32 * y2;
  const float x3 = y3 / D3;
  const float x2 = y2 / D2 - L32 * x3;
  const float x1 = y1 - L21 * x2 - L31 * x3;
  return v3(x1 * s1, x2 * s2, x3 * s3);
}

// ---------------------------------------------------------------------------
// records
// ---------------------------------------------------------------------------
// Stable Neo-Hookean tet: ids; rows of Dm^-1 with (6V signed, mu V, lambda V)
struct alignas(16) TetRec {
  U4 ids;
  F4 r0, r1, r2;
};
static_assert(sizeof(TetRec) == 64, "TetRec");

// One (body, tet) adjacency slot, stored at the slot's position in the
// static CSR so a body's slots are contiguous (coalesced lane loads).  With
// the shape-function gradients b_o of the three other vertices,
//   F = sum_o (x_o - x_self) b_o^T  and  w_self = -(b_1 + b_2 + b_3),
// so the stamp needs exactly one level of gathers (the three positions).
struct alignas(16) TetSlot {
  U4 other;   // other vertex ids; w = tet index
  F4 b1;      // w = mu V
  F4 b2;      // w = lambda V
  F4 b3;      // w = alpha_T = 1 + mu / lambda
};
static_assert(sizeof(TetSlot) == 64, "TetSlot");

struct alignas(16) ContactRec {
  F4 rA;       // w = feature key bits
  F4 rB;       // w = stick flag
  F4 C0;
  F4 lambda;
  F4 penalty;
};
struct alignas(16) ManifoldRec {
  U4 header;        // bodyA, bodyB, numContacts, flags (1 active, 2 A round, 4 B round)
  U4 colliderPair;  // colliderA, colliderB
  F4 basisN;        // w = dynamic friction
  F4 basisT1;       // w = static friction
  ContactRec c[kMaxContacts];
};
static_assert(sizeof(ManifoldRec) == 704, "ManifoldRec");

struct alignas(16) JointRec {
  U4 header;      // bodyA (kWorld = world), bodyB, broken, flags
  F4 rA;          // w = stiffnessLin
  F4 rB;          // w = stiffnessAng
  F4 C0Lin;       // w = torqueArm
  F4 C0Ang;       // w = fracture
  F4 lambdaLin, lambdaAng, penaltyLin, penaltyAng;
  F4 restRel;     // rest relative rotation
  F4 hingeAxis;   // w != 0 -> hinge
  F4 motor;       // target, effort limit, pad, kp
  F4 limits;      // lo, hi, kd, pad
  F4 dynamics;    // armature, predicted twist, explicit effort
};
static_assert(sizeof(JointRec) == 224, "JointRec");

struct alignas(16) SpringRec {
  U4 header;      // bodyA, bodyB, hard, pad
  F4 rA;          // w = stiffness / cap
  F4 rB;          // w = rest
  F4 dual;        // lambda, penalty, C0
};

// One (particle, element) slot of the cloth topology, stored at the slot's
// position in the static CSR (parallel to the entry list).  Partner ids carry
// kSnapBit when the partner shares the particle's static color.
//   membrane (StVK, 3 vertices):  u = {o1, o2, bits(mu A), bits(lambda A)},
//       f = {w1.x, w1.y, w2.x, w2.y}: shape-function gradients of the two
//       partners, F = sum_o (x_o - x_self) w_o^T, w_self = -(w1 + w2)
//   bend (Bergou hinge, 4 vertices): u = {o1, o2, o3, bits(s)},
//       f = {K1, K2, K3, K_self}: m = sum_o K_o (x_o - x_self)
//   particle spring: u = {other, spring index, hard, 0}, f = {k, rest, 0, 0}
struct alignas(16) ClothSlot {
  U4 u;
  F4 f;
};
static_assert(sizeof(ClothSlot) == 32, "ClothSlot");

// Unified element contact (V-T, rigid feature - T, E-E): 4-slot weighted
// stencil, C = C0 (1 - alpha) + sum_k w_k [n|t1|t2] . dq_k
struct alignas(16) SoftRec {
  U4 ids;        // participants; kWorld = unused slot
  F4 normal;     // xyz normal (pushes slot 0 along +n), w friction
  F4 anchorA;    // xyz rigid slot-0 anchor (rigid) or (restT, 0, 0); w = flag bits
  F4 weights;    // signed weight per slot
  F4 C0;         // xyz constraint at the step-start reference; w = lambda cap
  F4 lambda;     // xyz duals; w = persistence key A (bits)
  F4 penalty;    // xyz penalties; w = persistence key B (bits)
};
static_assert(sizeof(SoftRec) == 112, "SoftRec");
constexpr uint32_t kSoftRigidA = 1u, kSoftSideNeg = 2u, kSoftKindShift = 2, kSoftRoundA = 32u;
constexpr uint32_t kSoftVT = 1u, kSoftRT = 2u, kSoftEE = 3u;

// surface pair stream record: kind (0 V-T, 1 E-E), owner, other
constexpr uint32_t kDatVT = 0u, kDatEE = 1u;
constexpr uint32_t kElemEdgeBit = 0x80000000u;

// ---------------------------------------------------------------------------
// per-step parameters (passed by value as a kernel argument)
// ---------------------------------------------------------------------------
struct StepParams {
  float dt, gravity, alpha, betaLin, betaAng, gamma;
  float lambdaMax, maxSpeed, cellSize, collisionMargin;
  float particleDamping, rigidLinearDamping, rigidAngularDamping;
  uint32_t numBodies, numColliders, numJoints, numSprings, numTets;
  uint32_t mapCapacity, maxPairs, gridHashSize, numHashed, numGlobals, numHashedRigid;
  uint32_t numBpLane, numBpWave;  // pair producers on the lane / wave paths
  // 1 in the iterate-to-residual chunks: sweep kernels first read the
  // frame's end flag (CTR_CONV_DONE) and exit once it is set; 0 elsewhere,
  // where the flag is never set, so the kernels skip that dependent load
  uint32_t convCheck;
  uint32_t frictionCombineMode, numExclusions, iterations;
  uint32_t staticPalette;   // colors fixed for the scene's lifetime (tets / cloth present)
  // cloth surfaces (Planar-DAT pair stream + OGC element contacts)
  uint32_t numTris, numEdges, numParticles, elemHashSize, maxSoft, softMapCapacity, maxDatPairs;
  uint32_t numContactEdges;  // edges [0, numContactEdges) carry E-E contacts; later ones bound motion only
  uint32_t pairStream;       // the V-T / E-E pair stream and its contacts run
  uint32_t datOn;            // the Planar-DAT limiter passes run
  uint32_t isoOn;            // the isotropic OGC limiter and its tracked contacts run
  float elemCellSize, elemMargin, datQueryRadius, surfaceContactCellSize, clothViscosity;
  float fineCellSize;       // > 0: candidate search grid of 2 x the largest hashed radius
};

// ManifoldRec.header.w flag: both bodies dynamic and of the same static
// color -- the partner's pose is read from the start-of-iteration snapshot
constexpr uint32_t kManifoldSameColor = 0x100u;

AVBD_HD float combine_friction(float a, float b, uint32_t mode) {
  if (mode == 1u) return a * b;
  if (mode == 2u) return fminf(a, b);
  if (mode == 3u) return fmaxf(a, b);
  if (mode == 4u) return 0.5f * (a + b);
  return sqrtf(fmaxf(a * b, 0.0f));
}

}  // namespace avbd
