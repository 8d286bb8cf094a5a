// oracle.cpp -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
//
// A serial CPU restatement of the per-step AVBD solve of LuckyIYI/avbd-metal,
// templated on the scalar type so the same code runs in FP32 (parity) and
// FP64 (error reference).  Each function names the reference source it
// restates; paths are relative to the reference repository root.
//
// Semantics choices (the reference's CPU and GPU paths differ; this follows
// the GPU kernels, which are the spec for the step path):
//  * preconditioned, pivot-clamped 6x6 LDL^T (00_common.metal:137-221);
//  * trust region |dx| <= 0.35 r, |dth| <= 0.5 (40_solver.metal:1469-1489);
//  * mass-scaled contact dual cap (40_solver.metal:1820);
//  * sorted per-body force order: joints, springs, manifolds, tets
//    (40_solver.metal:220-243);
//  * a contact partner that shares the body's color is read from the
//    start-of-iteration pose snapshot (the backend closes the reference's
//    same-color Jacobi race this way; the oracle mirrors the rule).
// Covered: box / sphere / capsule / torus / convex-hull colliders, kinematic
// spinners, joints, springs, tets, cloth (Planar-DAT and the isotropic OGC
// limiter), the automatic limiter selection and the CPU path's step order.
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <set>
#include <vector>

// the spinner's per-frame rotation (x, y, z, w) by omega * dt about its unit
// axis, in float as the host computes it before every frame
static void oracle_spinner_rotation(const float* axisOmega, float dt, float* dq) {
  const float angle = axisOmega[3] * dt;
  const double half = 0.5 * (double)angle;
  const float sn = (float)std::sin(half), cs = (float)std::cos(half);
  dq[0] = axisOmega[0] * sn;
  dq[1] = axisOmega[1] * sn;
  dq[2] = axisOmega[2] * sn;
  dq[3] = cs;
}

namespace {

template <typename T>
struct V {
  T x, y, z;
  V() : x(0), y(0), z(0) {}
  V(T a, T b, T c) : x(a), y(b), z(c) {}
  V operator+(const V& o) const { return V(x + o.x, y + o.y, z + o.z); }
  V operator-(const V& o) const { return V(x - o.x, y - o.y, z - o.z); }
  V operator-() const { return V(-x, -y, -z); }
  V operator*(T s) const { return V(x * s, y * s, z * s); }
  V operator/(T s) const { return V(x / s, y / s, z / s); }
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> V<T> mulv(const V<T>& a, const V<T>& b) { return V<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <typename T> T dotv(const V<T>& a, const V<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> V<T> crossv(const V<T>& a, const V<T>& b) {
  return V<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> T lenv(const V<T>& a) { return std::sqrt(dotv(a, a)); }
template <typename T> V<T> normv(const V<T>& a) { return a * (T(1) / std::sqrt(dotv(a, a))); }

template <typename T>
struct Q {
  T x, y, z, w;
};
template <typename T> Q<T> qmul(const Q<T>& a, const Q<T>& b) {
  return Q<T>{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename T> Q<T> qconj(const Q<T>& q) { return Q<T>{-q.x, -q.y, -q.z, q.w}; }
template <typename T> Q<T> qinv(const Q<T>& q) {
  T n = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  return Q<T>{-q.x / n, -q.y / n, -q.z / n, q.w / n};
}
// 00_common.metal:98-101
template <typename T> V<T> qrot(const Q<T>& q, const V<T>& v) {
  V<T> u(q.x, q.y, q.z);
  V<T> t = crossv(u, v) * T(2);
  return v + t * q.w + crossv(u, t);
}
// 00_common.metal:113-118 (canonical hemisphere)
template <typename T> V<T> qsub(const Q<T>& a, const Q<T>& b) {
  Q<T> r = qmul(a, qinv(b));
  if (r.w < T(0)) r = Q<T>{-r.x, -r.y, -r.z, -r.w};
  return V<T>(r.x, r.y, r.z) * T(2);
}
// 00_common.metal:119-123
template <typename T> Q<T> qaddw(const Q<T>& q, const V<T>& w) {
  Q<T> d = qmul(Q<T>{w.x, w.y, w.z, T(0)}, q);
  Q<T> r{q.x + T(0.5) * d.x, q.y + T(0.5) * d.y, q.z + T(0.5) * d.z, q.w + T(0.5) * d.w};
  T s = T(1) / std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
  return Q<T>{r.x * s, r.y * s, r.z * s, r.w * s};
}
template <typename T> V<T> xformv(const V<T>& p, const Q<T>& q, const V<T>& v) { return qrot(q, v) + p; }

// row-major 3x3
template <typename T>
struct M {
  T a[3][3];
  static M zero() { M m; memset(m.a, 0, sizeof(m.a)); return m; }
  static M diag(const V<T>& d) { M m = zero(); m.a[0][0] = d.x; m.a[1][1] = d.y; m.a[2][2] = d.z; return m; }
  static M outer(const V<T>& u, const V<T>& v) {
    M m;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m.a[i][j] = u[i] * v[j];
    return m;
  }
  M operator+(const M& o) const { M m; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m.a[i][j] = a[i][j] + o.a[i][j]; return m; }
  M operator*(T s) const { M m; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m.a[i][j] = a[i][j] * s; return m; }
  M mul(const M& o) const {
    M m = zero();
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { T s = 0; for (int k = 0; k < 3; k++) s += a[i][k] * o.a[k][j]; m.a[i][j] = s; }
    return m;
  }
  V<T> mulv(const V<T>& v) const {
    return V<T>(a[0][0] * v.x + a[0][1] * v.y + a[0][2] * v.z, a[1][0] * v.x + a[1][1] * v.y + a[1][2] * v.z,
                a[2][0] * v.x + a[2][1] * v.y + a[2][2] * v.z);
  }
  M tr() const { M m; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m.a[i][j] = a[j][i]; return m; }
};
template <typename T> M<T> skew(const V<T>& r) {
  M<T> m = M<T>::zero();
  m.a[0][1] = -r.z; m.a[0][2] = r.y; m.a[1][0] = r.z; m.a[1][2] = -r.x; m.a[2][0] = -r.y; m.a[2][1] = r.x;
  return m;
}
// R diag(I) R^T (00_common.metal:102-110)
template <typename T> M<T> worldInertia(const Q<T>& q, const V<T>& I) {
  V<T> c0 = qrot(q, V<T>(1, 0, 0)), c1 = qrot(q, V<T>(0, 1, 0)), c2 = qrot(q, V<T>(0, 0, 1));
  return M<T>::outer(c0, c0) * I.x + M<T>::outer(c1, c1) * I.y + M<T>::outer(c2, c2) * I.z;
}

// 00_common.metal:137-221 -- Jacobi preconditioning, LDL^T, pivot clamp 1e-6.
template <typename T>
void solve6(const T A[6][6], const T b[6], T x[6]) {
  T s[6], P[6][6], rb[6];
  for (int i = 0; i < 6; i++) s[i] = T(1) / std::sqrt(std::max(A[i][i], T(1e-30)));
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) P[i][j] = (i == j) ? T(1) : A[i][j] * s[i] * s[j];
    rb[i] = b[i] * s[i];
  }
  T L[6][6] = {}, D[6];
  const T eps = T(1e-6);
  for (int j = 0; j < 6; j++) {
    T d = P[j][j];
    for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k] * D[k];
    D[j] = std::max(d, eps);
    for (int i = j + 1; i < 6; i++) {
      T v = P[i][j];
      for (int k = 0; k < j; k++) v -= L[j][k] * L[i][k] * D[k];
      L[i][j] = v / D[j];
    }
  }
  T y[6];
  for (int i = 0; i < 6; i++) {
    T v = rb[i];
    for (int k = 0; k < i; k++) v -= L[i][k] * y[k];
    y[i] = v;
  }
  T z[6];
  for (int i = 0; i < 6; i++) z[i] = y[i] / D[i];
  for (int i = 5; i >= 0; i--) {
    T v = z[i];
    for (int k = i + 1; k < 6; k++) v -= L[k][i] * x[k];
    x[i] = v;
  }
  for (int i = 0; i < 6; i++) x[i] *= s[i];
}

template <typename T> T clampT(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }
template <typename T> bool finiteT(T v) { return std::isfinite(v); }

const uint32_t WORLD = 0xFFFFFFFFu;
const uint32_t FKJ = 0, FKS = 1, FKM = 2, FKT = 3, FKMEM = 5, FKBEND = 6;

// ---------------------------------------------------------------------------
// records (GPUTypes.swift:80-173 field meanings)
// ---------------------------------------------------------------------------
template <typename T>
struct Contact {
  V<T> rA, rB, C0, lambda, penalty;
  uint32_t feature = 0;
  T stick = 0;
};
template <typename T>
struct Manifold {
  uint32_t a = 0, b = 0, n = 0, flags = 0, ca = 0, cb = 0;
  V<T> nrm, t1;
  T dynF = 0, statF = 0;
  Contact<T> c[8];
};
template <typename T>
struct Joint {
  float rate;  // legacy servo rate of an implicit-PD motor (0: none)
  uint32_t a, b, broken, flags;
  V<T> rA, rB;
  T stiffLin, stiffAng, torqueArm, fracture;
  V<T> C0Lin, C0Ang, lambdaLin, lambdaAng, penaltyLin, penaltyAng;
  V<T> penLin0, penAng0;
  Q<T> restRel;
  V<T> axis;
  bool hinge;
  T motorTarget, motorTorque, motorZ, kp, limLo, limHi, kd, armature, predTwist, effort;
};
template <typename T>
struct Spring {
  uint32_t a, b, hard;
  V<T> rA, rB;
  T stiffness, rest, lambda, penalty, C0;
};
template <typename T>
struct Tet {
  uint32_t ids[4];
  V<T> r[3];
  T six, muV, lamV;
};

// cloth: StVK membrane (2x2 inverse rest shape, column major), Bergou hinge,
// unified element contact (GPUTypes.swift:127-148)
template <typename T>
struct MembraneT { uint32_t ids[3]; T dm[4]; T muA, lamA; };
template <typename T>
struct BendT { uint32_t ids[4]; T K[4]; T s; };
template <typename T>
struct SoftT {
  uint32_t ids[4];
  V<T> n, anchor, lambda, penalty;
  T friction, C0, lamCap, w[4];
  uint32_t flags, keyA, keyB;
};

uint32_t cellHash32(int cx, int cy, int cz, uint32_t group, uint32_t hashSize) {
  // 20_broadphase.metal:16-28
  uint32_t h = ((uint32_t)cx * 92837111u) ^ ((uint32_t)cy * 689287499u) ^ ((uint32_t)cz * 283923481u);
  h ^= group * 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h & (hashSize - 1u);
}

uint32_t nextPow2(uint32_t v) { uint32_t n = 1; while (n < v) n <<= 1; return n; }

}  // namespace

// ---------------------------------------------------------------------------
// the solver
// ---------------------------------------------------------------------------
struct oracle_base {
  virtual ~oracle_base() {}
};

template <typename T>
struct Oracle : oracle_base {
  avbd_settings set{};
  uint32_t nb = 0, nc = 0;
  // bodies
  std::vector<V<T>> pos, vel, angVel, prevVel, initLin, inertLin, moment, size;
  // low part of the iterate: the position is pos + lo exactly (pos rounded);
  // zero at the frame start, carried by the primal update (TwoSum)
  std::vector<V<T>> lo;
  std::vector<Q<T>> rot, initAng, inertAng;
  std::vector<V<T>> snapLin;
  std::vector<Q<T>> snapAng;
  std::vector<T> mass, radiusSigned, gravityScale, friction;
  std::vector<uint8_t> particle;
  std::vector<uint32_t> clothGroup;
  // colliders
  std::vector<uint32_t> cOwner, cGroup, cShared, cFlags;
  std::vector<std::vector<V<T>>> cHull;  // convex hull vertices per collider (collider-local), empty = primitive
  std::vector<V<T>> cSize, cLocal;
  std::vector<Q<T>> cLocalRot;
  std::vector<T> cRad, cStatF, cDynF;
  // kinematic spinners (GPUSolver.swift:628-637): spinner k turns body
  // spinBody[k] about a unit axis at omega rad/s; spinW / spinOn are the
  // per-body angular velocity the contacts' surface shift reads
  std::vector<uint32_t> spinBody;
  std::vector<std::array<float, 4>> spinAxisOmega;
  std::vector<V<T>> spinW;
  std::vector<uint8_t> spinOn;
  std::vector<uint32_t> hashed, globals;
  uint32_t nHashedRigid = 0;
  float cellSize = 0.1f;
  uint32_t hashSize = 64, maxPairs = 4096;
  std::vector<std::pair<uint32_t, uint32_t>> excl;
  // elements
  std::vector<Joint<T>> joints;
  std::vector<Spring<T>> springs;
  std::vector<Tet<T>> tets;
  std::vector<std::vector<uint32_t>> staticAdj;  // per body, sorted entries (joints, springs, tets)
  // cloth surfaces (GPUSolver.swift:1289-1690, 25_cloth.metal, 27_planar_dat.metal)
  using Membrane = MembraneT<T>;
  using Bend = BendT<T>;
  using Soft = SoftT<T>;
  uint32_t nTris = 0, numContactEdges = 0;
  bool pairStream = false;  // the V-T / E-E pair stream and its contacts (with the Planar-DAT limiter)
  bool datOn = false;       // the Planar-DAT limiter: a self-colliding (authored) surface exists
  // isotropic OGC limiter (the reference's automatic choice where tets exist,
  // GPUSolver.swift:93-102): tracked V-T / E-E contacts and per-vertex bounds
  bool isoOn = false;
  std::vector<std::array<uint32_t, 4>> triAdj;         // up to 3 edge-adjacent collision triangles
  std::vector<uint32_t> nbr2Start, nbr2Count, nbr2List;  // sorted 2-ring per vertex
  std::vector<uint32_t> vEdgeStart, vEdgeCount, vEdgeList;  // incident edges per vertex
  std::vector<uint8_t> clothVert;                        // vertex of an authored triangle
  std::vector<std::array<uint32_t, 4>> vtTrack, vtTrackPrev, eeTrack, eeTrackPrev;
  std::vector<T> ogcBound;                               // squared clearance bound per body
  std::vector<V<T>> ogcPrev;                             // the pose the bounds hold around
  uint64_t ogcExceeded = 0;                              // primal truncations since the last refresh check
  uint32_t ogcTruncFrame = 0, ogcRefreshFrame = 0;       // this frame's truncations / refreshes
  std::vector<std::array<uint32_t, 3>> tris;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  std::vector<uint32_t> nbrStart, nbrCount, nbrList, particles, selfColl, hashedRigidList;
  std::vector<Membrane> membranes;
  std::vector<Bend> bends;
  float elemMargin = 0, datQueryRadius = 0, elemCellSize = 0, surfaceContactCellSize = 0;
  uint32_t elemHashSize = 64;
  uint32_t maxDatPairs = 1, maxSoft = 1;
  std::vector<std::array<uint32_t, 3>> datPairs;
  uint32_t datTotal = 0, datVT = 0, datOverflow = 0, datNonfinite = 0, softCandidates = 0, softRT = 0;
  // Planar-DAT truncation state: fixed-plane anchors, per-vertex step
  // fractions, and the frame's counters (peak pair demand over both queries)
  std::vector<V<T>> datAnchor;
  std::vector<T> datT;
  uint32_t datPeak = 0, datPeakVT = 0, datInvalid = 0, datTrunc = 0, datVTDeg = 0, datEEDeg = 0;
  std::vector<Soft> soft, prevSoft;
  std::map<std::pair<uint32_t, uint32_t>, int> softMap;  // keys of prevSoft -> index
  std::vector<std::vector<uint32_t>> softAdj;
  // frame state
  std::vector<Manifold<T>> manifolds, prevManifolds;
  std::map<std::pair<uint32_t, uint32_t>, int> prevMap;  // collider pair -> index in prevManifolds
  std::vector<uint32_t> cellStart, cellCount, cellBodies;
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  uint32_t pairCandidates = 0, activeManifolds = 0, colorConflicts = 0;
  // colors
  bool dynamicColoring = false;
  std::vector<uint32_t> colors;
  std::vector<std::vector<uint32_t>> buckets;
  int64_t frame = 0;

  bool dyn(uint32_t b) const { return b != WORLD && mass[b] > T(0); }

  // -------------------------------------------------------------------------
  // upload (GPUSolver.swift:901-2081)
  // -------------------------------------------------------------------------
  int init(const avbd_scene_desc* d, int maxPairsPerBody) {
    set = d->settings;
    nb = d->num_bodies;
    nc = d->num_colliders;
    auto r3 = [](const float* p, uint32_t i) { return V<T>(p[3 * i], p[3 * i + 1], p[3 * i + 2]); };
    auto r4 = [](const float* p, uint32_t i) { return Q<T>{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; };
    pos.resize(nb); lo.assign(nb, V<T>()); vel.resize(nb); angVel.resize(nb); prevVel.resize(nb); initLin.resize(nb); inertLin.resize(nb);
    moment.resize(nb); size.resize(nb); rot.resize(nb); initAng.resize(nb); inertAng.resize(nb);
    mass.resize(nb); radiusSigned.resize(nb); gravityScale.resize(nb); friction.resize(nb); particle.resize(nb);
    const float pi = 3.14159265358979f;
    for (uint32_t i = 0; i < nb; i++) {
      // mass / inertia per shape: computed in float as the upload does (GPUSolver.swift:914-962)
      float sx = d->body_size[3 * i], sy = d->body_size[3 * i + 1], sz = d->body_size[3 * i + 2];
      uint32_t sh = d->body_shape[i];
      if (sh == AVBD_SHAPE_SPHERE) { sy = sx; sz = sx; }
      float den = d->body_density[i], m = 0, r = 0, ix = 0, iy = 0, iz = 0;
      if (sh == AVBD_SHAPE_BOX) {
        m = den > 0 ? sx * sy * sz * den : 0.0f;
        ix = (sy * sy + sz * sz) / 12.0f * m; iy = (sx * sx + sz * sz) / 12.0f * m; iz = (sx * sx + sy * sy) / 12.0f * m;
        float hx = sx * 0.5f, hy = sy * 0.5f, hz = sz * 0.5f;
        r = std::sqrt(hx * hx + hy * hy + hz * hz);
      } else if (sh == AVBD_SHAPE_SPHERE) {
        float rr = sx / 2.0f;
        m = den > 0 ? 4.0f / 3.0f * pi * rr * rr * rr * den : 0.0f;
        ix = iy = iz = 0.4f * m * rr * rr;
        r = rr;
      } else if (sh == AVBD_SHAPE_CAPSULE) {
        // GPUSolver.swift:940-945: cylinder of length L plus two hemispheres
        float L = sx, rr = sy;
        m = den > 0 ? pi * rr * rr * (L + 4.0f * rr / 3.0f) * den : 0.0f;
        float iAxis = 0.5f * m * rr * rr, iPerp = m * (L * L / 12.0f + rr * rr / 4.0f);
        ix = iy = iPerp;
        iz = iAxis;
        r = L / 2.0f + rr;
      } else if (sh == AVBD_SHAPE_TORUS) {
        // GPUSolver.swift:933-939: solid torus, spine radius R, tube radius r
        float R = sx, rr = sy;
        m = den > 0 ? 2.0f * pi * pi * R * rr * rr * den : 0.0f;
        float iDia = m * (R * R / 2.0f + 5.0f * rr * rr / 8.0f), iAxis = m * (R * R + 3.0f * rr * rr / 4.0f);
        ix = iy = iDia;
        iz = iAxis;
        r = R + rr;
      } else {
        return 4;
      }
      if (d->body_mass && !std::isnan(d->body_mass[i])) {
        m = d->body_mass[i];
        ix = d->body_inertia[3 * i]; iy = d->body_inertia[3 * i + 1]; iz = d->body_inertia[3 * i + 2];
      }
      particle[i] = d->body_is_particle ? d->body_is_particle[i] : 0;
      mass[i] = m;
      moment[i] = V<T>(ix, iy, iz);
      size[i] = V<T>(sx, sy, sz);
      radiusSigned[i] = particle[i] ? -T(r) : T(r);
      pos[i] = r3(d->body_position, i);
      rot[i] = d->body_rotation ? r4(d->body_rotation, i) : Q<T>{0, 0, 0, 1};
      vel[i] = d->body_velocity ? r3(d->body_velocity, i) : V<T>();
      prevVel[i] = vel[i];
      angVel[i] = V<T>();
      gravityScale[i] = d->body_gravity_scale ? d->body_gravity_scale[i] : 1.0f;
      friction[i] = d->body_friction[i];
    }
    // colliders (GPUSolver.swift:971-1090)
    cOwner.resize(nc); cGroup.resize(nc); cShared.resize(nc); cFlags.resize(nc); cSize.resize(nc); cLocal.resize(nc);
    cHull.assign(nc, {});
    uint32_t hullBase = 0;
    cLocalRot.resize(nc); cRad.resize(nc); cStatF.resize(nc); cDynF.resize(nc);
    std::vector<float> radii;
    uint32_t enabled = 0;
    bool isolated = false;
    std::vector<float> rad32(nc);
    for (uint32_t i = 0; i < nc; i++) {
      uint32_t b = d->collider_body[i], sh = d->collider_shape[i];
      float sx = d->collider_size[3 * i], sy = d->collider_size[3 * i + 1], sz = d->collider_size[3 * i + 2];
      if (sh == AVBD_SHAPE_SPHERE) { sy = sx; sz = sx; }
      float r;
      if (sh == AVBD_SHAPE_SPHERE) r = sx / 2.0f;
      else if (sh == AVBD_SHAPE_CAPSULE) r = sx / 2.0f + sy;
      else if (sh == AVBD_SHAPE_TORUS) r = sx + sy;
      else { float hx = sx * 0.5f, hy = sy * 0.5f, hz = sz * 0.5f; r = std::sqrt(hx * hx + hy * hy + hz * hz); }
      const uint32_t hn = d->collider_hull_count ? d->collider_hull_count[i] : 0u;
      if (hn) {  // GPUSolver.swift:994-997: the farthest vertex bounds the hull
        r = 0.0f;
        for (uint32_t k = 0; k < hn; k++) {
          const float* v = d->hull_vertices + 3 * (hullBase + k);
          cHull[i].push_back(V<T>(v[0], v[1], v[2]));
          r = std::max(r, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        }
        hullBase += hn;
      }
      rad32[i] = r;
      bool en = !d->collider_enabled || d->collider_enabled[i];
      cOwner[i] = b;
      cGroup[i] = d->collider_group ? d->collider_group[i] : 0;
      cShared[i] = (!d->collider_shared || d->collider_shared[i]) ? 1 : 0;
      cSize[i] = V<T>(sx, sy, sz);
      cRad[i] = particle[b] ? -T(r) : T(r);
      cLocal[i] = d->collider_local_position ? r3(d->collider_local_position, i) : V<T>();
      Q<T> lq = d->collider_local_rotation ? r4(d->collider_local_rotation, i) : Q<T>{0, 0, 0, 1};
      T qn = std::sqrt(lq.x * lq.x + lq.y * lq.y + lq.z * lq.z + lq.w * lq.w);
      cLocalRot[i] = Q<T>{lq.x / qn, lq.y / qn, lq.z / qn, lq.w / qn};
      cStatF[i] = d->collider_friction[i];
      cDynF[i] = d->collider_dynamic_friction ? d->collider_dynamic_friction[i] : d->collider_friction[i];
      uint32_t f = hn ? 4u : sh;
      if (particle[b]) f |= 0x10u;
      if (d->collider_world_round_anchor && d->collider_world_round_anchor[i]) f |= 0x20u;
      cFlags[i] = f;
      if (en) {
        enabled++;
        if (cGroup[i] != 0) isolated = true;
        bool bodyDyn = (d->body_mass && !std::isnan(d->body_mass[b])) ? d->body_mass[b] > 0 : d->body_density[b] > 0;
        if (bodyDyn) radii.push_back(r);
      }
    }
    std::sort(radii.begin(), radii.end());
    float median = radii.empty() ? 0.5f : radii[radii.size() / 2];
    float thr = median * 4.0f;
    for (uint32_t i = 0; i < nc; i++) {
      bool en = !d->collider_enabled || d->collider_enabled[i];
      if (!en) continue;
      if (rad32[i] > thr || (isolated && cGroup[i] == 0)) globals.push_back(i);
      else hashed.push_back(i);
    }
    float maxR = 0.05f;
    for (uint32_t i : hashed) maxR = std::max(maxR, rad32[i]);
    cellSize = maxR * 2.0f;
    for (uint32_t i : hashed)
      if (!particle[cOwner[i]]) { nHashedRigid++; hashedRigidList.push_back(i); }
    maxPairs = std::max<uint32_t>(4096u, enabled * (uint32_t)maxPairsPerBody);
    hashSize = nextPow2(std::max<uint32_t>(64u, 2u * std::max<uint32_t>(1u, enabled)));
    // spinners: the last spinner of a body sets its surface angular velocity
    spinW.assign(nb, V<T>());
    spinOn.assign(nb, 0);
    for (uint32_t k = 0; k < d->num_spinners; k++) {
      const uint32_t b = d->spinner_body[k];
      if (b >= nb) return 5;
      const float* a = d->spinner_axis_omega + 4 * k;
      spinBody.push_back(b);
      spinAxisOmega.push_back({a[0], a[1], a[2], a[3]});
      spinW[b] = V<T>(a[0] * a[3], a[1] * a[3], a[2] * a[3]);
      spinOn[b] = 1;
    }
    // joints (GPUSolver.swift:1092-1193)
    for (uint32_t i = 0; i < d->num_joints; i++) {
      Joint<T> j{};
      int32_t a = d->joint_body_a[i], b = d->joint_body_b[i];
      float kl = d->joint_stiffness_lin[i], ka = d->joint_stiffness_ang[i];
      float fr = d->joint_fracture ? d->joint_fracture[i] : INFINITY;
      j.a = a >= 0 ? (uint32_t)a : WORLD;
      j.b = (uint32_t)b;
      j.broken = 0;
      j.flags = (std::isinf(kl) ? 1u : 0u) | (std::isinf(ka) ? 2u : 0u) | (std::isfinite(fr) ? 4u : 0u) |
                ((d->joint_fracture_linear && d->joint_fracture_linear[i]) ? 8u : 0u);
      j.rA = r3(d->joint_ra, i);
      j.rB = r3(d->joint_rb, i);
      j.stiffLin = std::min(kl, 3.0e10f);
      j.stiffAng = std::min(ka, 3.0e10f);
      V<T> sA = a >= 0 ? size[a] : V<T>();
      V<T> ss = sA + size[b];
      j.torqueArm = dotv(ss, ss);
      j.fracture = std::min(fr, 3.0e18f);
      Q<T> qa = a >= 0 ? rot[a] : Q<T>{0, 0, 0, 1};
      Q<T> rel = qmul(qinv(qa), rot[b]);
      T rn = std::sqrt(rel.x * rel.x + rel.y * rel.y + rel.z * rel.z + rel.w * rel.w);
      j.restRel = Q<T>{rel.x / rn, rel.y / rn, rel.z / rn, rel.w / rn};
      j.hinge = d->joint_hinge_axis && !std::isnan(d->joint_hinge_axis[3 * i]);
      const float* mo = d->joint_motor ? d->joint_motor + 6 * i : nullptr;
      j.motorTorque = 0; j.motorTarget = 0; j.kp = 0; j.kd = 0; j.armature = 0; j.limLo = 1; j.limHi = -1;
      float lo = d->joint_limits ? d->joint_limits[2 * i] : 1.0f, hi = d->joint_limits ? d->joint_limits[2 * i + 1] : -1.0f;
      if (j.hinge) {
        j.axis = normv(r3(d->joint_hinge_axis, i));
        j.armature = mo ? mo[4] : 0.0f;
        if (mo && mo[1] > 0) {
          j.motorTarget = mo[0]; j.motorTorque = mo[1]; j.kp = mo[2]; j.kd = mo[3];
          j.rate = d->joint_motor_rate ? d->joint_motor_rate[i] : 0.0f;
          j.limLo = lo; j.limHi = hi;
          int mode = (int)mo[5];
          j.flags |= (mode == 1 ? (2u << 4) : (mode == 2 ? (3u << 4) : (1u << 4)));
        }
        if (lo < hi && !(mo && mo[1] > 0)) { j.limLo = lo; j.limHi = hi; j.kd = 0; }
      }
      T mA = a >= 0 ? mass[a] : T(0), mB = mass[b];
      T eff = 0;
      bool any = false;
      if (mA > 0) { eff = mA; any = true; }
      if (mB > 0) { eff = any ? std::min(eff, mB) : mB; any = true; }
      T dt = set.dt;
      T floorK = std::min(T(1.0e9), std::max(T(1), eff / std::max(dt * dt, T(1e-12))));
      if (std::isinf(kl)) j.penaltyLin = V<T>(floorK, floorK, floorK);
      if (std::isinf(ka)) j.penaltyAng = V<T>(floorK, floorK, floorK);
      if (kl > 0 && std::isfinite(kl)) { T k = std::min(kl, 1e9f); j.penaltyLin = V<T>(k, k, k); }
      if (ka > 0 && std::isfinite(ka)) { T k = std::min(ka, 1e9f); j.penaltyAng = V<T>(k, k, k); }
      j.penLin0 = j.penaltyLin;
      j.penAng0 = j.penaltyAng;
      joints.push_back(j);
    }
    // springs (GPUSolver.swift:1195-1212)
    for (uint32_t i = 0; i < d->num_springs; i++) {
      Spring<T> s{};
      s.a = d->spring_body_a[i];
      s.b = d->spring_body_b[i];
      s.hard = (d->spring_hard && d->spring_hard[i]) ? 1 : 0;
      s.rA = r3(d->spring_ra, i);
      s.rB = r3(d->spring_rb, i);
      s.stiffness = d->spring_stiffness[i];
      T rest = d->spring_rest[i];
      if (rest < 0) rest = lenv((pos[s.a] + qrot(rot[s.a], s.rA)) - (pos[s.b] + qrot(rot[s.b], s.rB)));
      s.rest = rest;
      springs.push_back(s);
    }
    // tets: Dm from spawn positions, rows of Dm^-1, 6V, mu V, lambda V (GPUSolver.swift:1214-1236)
    for (uint32_t t = 0; t < d->num_tets; t++) {
      Tet<T> g{};
      for (int k = 0; k < 4; k++) g.ids[k] = d->tet_ids[4 * t + k];
      double D[3][3];
      for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++)
          D[k][c] = (double)d->body_position[3 * g.ids[c + 1] + k] - (double)d->body_position[3 * g.ids[0] + k];
      double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) - D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                   D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
      double C[3][3];  // cofactor transpose / det = inverse
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
          C[i][j] = (D[i1][j1] * D[i2][j2] - D[i1][j2] * D[i2][j1]) / det;
        }
      for (int r = 0; r < 3; r++) g.r[r] = V<T>((T)(float)C[r][0], (T)(float)C[r][1], (T)(float)C[r][2]);
      double vol = std::fabs(det) / 6.0;
      g.six = (T)(float)det;
      g.muV = (T)(float)(d->tet_mu[t] * vol);
      g.lamV = (T)(float)(d->tet_lambda[t] * vol);
      tets.push_back(g);
    }
    // exclusions (GPUSolver.swift:1265-1288)
    std::set<std::pair<uint32_t, uint32_t>> ex;
    for (auto& j : joints) if (j.a != WORLD) ex.insert({std::min(j.a, j.b), std::max(j.a, j.b)});
    for (auto& s : springs) ex.insert({std::min(s.a, s.b), std::max(s.a, s.b)});
    for (uint32_t i = 0; i < d->num_exclusions; i++) {
      uint32_t a = d->exclusion_pairs[2 * i], b = d->exclusion_pairs[2 * i + 1];
      ex.insert({std::min(a, b), std::max(a, b)});
    }
    excl.assign(ex.begin(), ex.end());
    // collision surfaces: authored triangles and tet boundary faces; soft
    // collision groups over them (GPUSolver.swift:1476-1522)
    clothGroup.assign(nb, 0);
    if (d->num_tris > 0 || !tets.empty()) clothInit(d);
    // static force adjacency, sorted: joints, springs, tets
    staticAdj.assign(nb, {});
    for (uint32_t i = 0; i < joints.size(); i++) {
      const Joint<T>& j = joints[i];
      if (j.stiffLin == 0 && j.stiffAng == 0 && j.motorTorque == 0) continue;
      if (dyn(j.a)) staticAdj[j.a].push_back((FKJ << 28) | i);
      if (dyn(j.b)) staticAdj[j.b].push_back((FKJ << 28) | i);
    }
    for (uint32_t i = 0; i < springs.size(); i++) {
      if (dyn(springs[i].a)) staticAdj[springs[i].a].push_back((FKS << 28) | i);
      if (dyn(springs[i].b)) staticAdj[springs[i].b].push_back((FKS << 28) | i);
    }
    for (uint32_t t = 0; t < tets.size(); t++)
      for (int k = 0; k < 4; k++) if (dyn(tets[t].ids[k])) staticAdj[tets[t].ids[k]].push_back((FKT << 28) | t);
    for (uint32_t m = 0; m < membranes.size(); m++)
      for (int k = 0; k < 3; k++) if (dyn(membranes[m].ids[k])) staticAdj[membranes[m].ids[k]].push_back((FKMEM << 28) | m);
    for (uint32_t m = 0; m < bends.size(); m++)
      for (int k = 0; k < 4; k++) if (dyn(bends[m].ids[k])) staticAdj[bends[m].ids[k]].push_back((FKBEND << 28) | m);
    dynamicColoring = d->num_tets == 0 && d->num_tris == 0;
    // static greedy palette in body-index order (GPUSolver.swift:1956-2050)
    std::vector<std::set<uint32_t>> adj(nb);
    auto link = [&](uint32_t a, uint32_t b) {
      if (a == WORLD || b == WORLD) return;
      if (dyn(a) && dyn(b)) { adj[a].insert(b); adj[b].insert(a); }
    };
    for (auto& j : joints) if (j.stiffLin != 0 || j.stiffAng != 0 || j.motorTorque != 0) link(j.a, j.b);
    for (auto& s : springs) link(s.a, s.b);
    for (auto& t : tets) for (int i = 0; i < 4; i++) for (int k = i + 1; k < 4; k++) link(t.ids[i], t.ids[k]);
    colors.assign(nb, 0);
    for (uint32_t v = 0; v < nb; v++) {
      if (!dyn(v)) continue;
      uint64_t used = 0;
      for (uint32_t n : adj[v]) if (n < v && colors[n] < 64) used |= 1ull << colors[n];
      if (used == ~0ull) return 5;  // staticColorCapacity
      uint32_t c = 0;
      while (used & (1ull << c)) c++;
      colors[v] = c;
    }
    rebuildBuckets();
    snapLin = pos;
    snapAng = rot;
    return 0;
  }

  // Surface topology, membranes, hinges, grid / query parameters and
  // capacities (GPUSolver.swift:391-443, 1289-1690).  Element constants are
  // derived in double and rounded to float, as the records store them.
  // Boundary faces of the tet meshes: faces of exactly one tet, wound so the
  // normal points away from that tet's fourth vertex, in ascending
  // (a, b, c) order (GPUSolver.swift:872-897).
  static std::vector<std::array<uint32_t, 3>> tetBoundaryFaces(const avbd_scene_desc* d) {
    std::map<std::array<uint32_t, 3>, std::pair<std::array<uint32_t, 4>, int>> seen;  // sorted -> (tri + opposite, uses)
    const int fv[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    for (uint32_t t = 0; t < d->num_tets; t++) {
      const uint32_t* q = d->tet_ids + 4 * t;
      for (auto& f : fv) {
        std::array<uint32_t, 3> key{q[f[0]], q[f[1]], q[f[2]]};
        std::sort(key.begin(), key.end());
        auto it = seen.find(key);
        if (it == seen.end()) seen[key] = {{q[f[0]], q[f[1]], q[f[2]], q[f[3]]}, 1};
        else it->second.second++;
      }
    }
    auto P = [&](uint32_t i) {
      return V<float>(d->body_position[3 * i], d->body_position[3 * i + 1], d->body_position[3 * i + 2]);
    };
    std::vector<std::array<uint32_t, 3>> out;
    for (auto& kv : seen) {
      if (kv.second.second != 1) continue;
      auto f = kv.second.first;
      V<float> n = crossv(P(f[1]) - P(f[0]), P(f[2]) - P(f[0]));
      if (dotv(n, P(f[3]) - P(f[0])) > 0.0f) std::swap(f[1], f[2]);
      out.push_back({f[0], f[1], f[2]});
    }
    std::sort(out.begin(), out.end());
    return out;
  }

  void clothInit(const avbd_scene_desc* d) {
    // collision triangles: the authored ones, then the tet boundary faces;
    // membranes and hinges from the authored ones only (GPUSolver.swift:1289-1334)
    const uint32_t nAuthored = d->num_tris;
    std::vector<std::array<uint32_t, 3>> coll;
    for (uint32_t t = 0; t < nAuthored; t++) coll.push_back({d->tri_ids[3 * t], d->tri_ids[3 * t + 1], d->tri_ids[3 * t + 2]});
    for (const auto& f : tetBoundaryFaces(d)) coll.push_back(f);
    nTris = (uint32_t)coll.size();
    auto sizeX = [&](uint32_t i) { return d->body_size[3 * i]; };
    auto bp = [&](uint32_t i, int k) { return d->body_position[3 * i + k]; };
    auto fdist = [&](uint32_t a, uint32_t b) {
      float dx = bp(a, 0) - bp(b, 0), dy = bp(a, 1) - bp(b, 1), dz = bp(a, 2) - bp(b, 2);
      return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    std::set<uint64_t> seen;
    float maxElemR = 0.0f, maxPartR = 0.0f;
    for (uint32_t t = 0; t < nTris; t++) {
      // edges of the authored triangles come first: the E-E contact prefix
      if (t == nAuthored) numContactEdges = (uint32_t)edges.size();
      std::array<uint32_t, 3> id = coll[t];
      tris.push_back(id);
      uint32_t ev[3][2] = {{id[0], id[1]}, {id[1], id[2]}, {id[0], id[2]}};
      for (auto& e : ev) {
        uint64_t k = (uint64_t)std::min(e[0], e[1]) << 32 | std::max(e[0], e[1]);
        if (seen.insert(k).second) edges.push_back({std::min(e[0], e[1]), std::max(e[0], e[1])});
      }
      float m[3];
      for (int k = 0; k < 3; k++) m[k] = (bp(id[0], k) + bp(id[1], k) + bp(id[2], k)) / 3.0f;
      auto dm = [&](uint32_t v) {
        float dx = m[0] - bp(v, 0), dy = m[1] - bp(v, 1), dz = m[2] - bp(v, 2);
        return std::sqrt(dx * dx + dy * dy + dz * dz);
      };
      float thick = std::max(sizeX(id[0]), std::max(sizeX(id[1]), sizeX(id[2]))) / 2.0f;
      maxPartR = std::max(maxPartR, thick);
      maxElemR = std::max(maxElemR, std::max(dm(id[0]), std::max(dm(id[1]), dm(id[2]))) + thick);
    }
    if (nTris == nAuthored) numContactEdges = (uint32_t)edges.size();
    std::vector<std::vector<uint32_t>> nb1(nb);
    float edgeLenSum = 0.0f;
    for (auto& e : edges) {
      nb1[e.first].push_back(e.second);
      nb1[e.second].push_back(e.first);
      maxElemR = std::max(maxElemR, fdist(e.first, e.second) / 2.0f + std::max(sizeX(e.first), sizeX(e.second)) / 2.0f);
      maxPartR = std::max(maxPartR, std::max(sizeX(e.first), sizeX(e.second)) / 2.0f);
      edgeLenSum += fdist(e.first, e.second);
    }
    nbrStart.assign(nb, 0);
    nbrCount.assign(nb, 0);
    for (uint32_t i = 0; i < nb; i++) {
      std::sort(nb1[i].begin(), nb1[i].end());
      nbrStart[i] = (uint32_t)nbrList.size();
      nbrCount[i] = (uint32_t)nb1[i].size();
      nbrList.insert(nbrList.end(), nb1[i].begin(), nb1[i].end());
    }
    // connected surfaces, numbered by their first triangle; a surface with an
    // authored triangle self-collides, a tet boundary does not (GPUSolver.swift:1476-1522)
    std::vector<uint32_t> parent(nb);
    for (uint32_t i = 0; i < nb; i++) parent[i] = i;
    std::function<uint32_t(uint32_t)> find = [&](uint32_t x) { return parent[x] == x ? x : (parent[x] = find(parent[x])); };
    for (auto& t : tris) { uint32_t r = find(t[0]); parent[find(t[1])] = r; parent[find(t[2])] = r; }
    selfColl.assign(nb, 0);
    std::set<uint32_t> selfRoots;
    for (uint32_t t = 0; t < nAuthored; t++) selfRoots.insert(find(tris[t][0]));
    std::map<uint32_t, uint32_t> gid;
    std::vector<uint8_t> surf(nb, 0);
    for (auto& t : tris) {
      uint32_t r = find(t[0]);
      if (!gid.count(r)) { uint32_t nid = (uint32_t)gid.size() + 1; gid[r] = nid; }
      for (uint32_t v : t)
        if (particle[v]) { clothGroup[v] = gid[r]; selfColl[v] = selfRoots.count(r) ? 1 : 0; surf[v] = 1; }
    }
    datOn = !selfRoots.empty();
    pairStream = datOn;  // the V-T / E-E pair stream runs with the Planar-DAT limiter
    for (uint32_t i = 0; i < nb; i++) if (surf[i]) particles.push_back(i);
    // surface limiter (GPUSolver.swift:93-102): Planar-DAT for authored
    // (self-colliding) surfaces without tets, the isotropic OGC bound
    // wherever tets exist (the scene format carries no volumetric
    // self-collision flag, so tet bodies never self-collide); the isotropic
    // mode emits its own tracked contacts.
    // With one closed surface and no authored triangle every candidate pair
    // of the isotropic trackers is a same-solid pair, which they skip: no
    // bound, no contact, no truncation -- the mode is then left off, an
    // identity on the results.
    isoOn = d->num_tets > 0 && !particles.empty() && (gid.size() >= 2 || nAuthored > 0);
    if (isoOn) {
      datOn = false;
      pairStream = false;
    }
    clothVert.assign(nb, 0);
    for (uint32_t t = 0; t < nAuthored; t++)
      for (uint32_t v : tris[t]) clothVert[v] = 1;
    datAnchor.assign(nb, V<T>());
    datT.assign(nb, T(1));
    // StVK membranes: rest 2x2 in the triangle frame, inverse column major
    std::map<uint64_t, std::vector<std::pair<uint32_t, float>>> edgeTris;
    for (uint32_t t = 0; t < nAuthored; t++) {
      float mu = d->tri_mu ? d->tri_mu[t] : 0.0f;
      if (!(mu > 0.0f)) continue;
      float lam = d->tri_lambda ? d->tri_lambda[t] : mu, bend = d->tri_bend ? d->tri_bend[t] : 0.0f;
      uint32_t i0 = tris[t][0], i1 = tris[t][1], i2 = tris[t][2];
      V<double> x0(bp(i0, 0), bp(i0, 1), bp(i0, 2));
      V<double> d1 = V<double>(bp(i1, 0), bp(i1, 1), bp(i1, 2)) - x0, d2 = V<double>(bp(i2, 0), bp(i2, 1), bp(i2, 2)) - x0;
      V<double> n = crossv(d1, d2);
      double area = lenv(n) / 2.0;
      if (!(area > 1e-10)) continue;
      V<double> e1 = normv(d1), e2 = normv(crossv(normv(n), e1));
      double a = dotv(d1, e1), b = dotv(d1, e2), c = dotv(d2, e1), dd = dotv(d2, e2);
      double det = a * dd - b * c;
      if (!(std::fabs(det) > 1e-12)) continue;
      Membrane mb;
      mb.ids[0] = i0; mb.ids[1] = i1; mb.ids[2] = i2;
      mb.dm[0] = (T)(float)(dd / det); mb.dm[1] = (T)(float)(-b / det);
      mb.dm[2] = (T)(float)(-c / det); mb.dm[3] = (T)(float)(a / det);
      mb.muA = (T)(float)(mu * area);
      mb.lamA = (T)(float)(lam * area);
      membranes.push_back(mb);
      if (bend > 0.0f) {
        uint32_t ev[3][3] = {{i0, i1, i2}, {i1, i2, i0}, {i0, i2, i1}};
        for (auto& e : ev) edgeTris[(uint64_t)std::min(e[0], e[1]) << 32 | std::max(e[0], e[1])].push_back({e[2], bend});
      }
    }
    // Bergou hinges on edges shared by two bending triangles (edge-key order)
    for (auto& kv : edgeTris) {
      if (kv.second.size() != 2) continue;
      uint32_t v0 = (uint32_t)(kv.first >> 32), v1 = (uint32_t)(kv.first & 0xFFFFFFFFu);
      uint32_t o2 = kv.second[0].first, o3 = kv.second[1].first;
      V<double> p0(bp(v0, 0), bp(v0, 1), bp(v0, 2)), p1(bp(v1, 0), bp(v1, 1), bp(v1, 2));
      double L = lenv(p0 - p1);
      if (!(L > 1e-9)) continue;
      double q[4][2] = {{0, 0}, {L, 0}, {0, 0}, {0, 0}};
      bool ok = true;
      for (int f = 0; f < 2 && ok; f++) {
        uint32_t o = f == 0 ? o2 : o3;
        V<double> po(bp(o, 0), bp(o, 1), bp(o, 2));
        double l0 = lenv(po - p0), l1 = lenv(po - p1);
        double x = (L * L + l0 * l0 - l1 * l1) / (2 * L), h2 = l0 * l0 - x * x;
        if (!(h2 > 1e-12)) { ok = false; break; }
        q[2 + f][0] = x;
        q[2 + f][1] = (f == 0 ? 1.0 : -1.0) * std::sqrt(h2);
      }
      if (!ok) continue;
      // K = generalized 4-D cross product of the rows {1}, {x}, {y}
      auto minor = [&](int c0, int c1, int c2) {
        double m00 = 1, m01 = 1, m02 = 1;
        double m10 = q[c0][0], m11 = q[c1][0], m12 = q[c2][0];
        double m20 = q[c0][1], m21 = q[c1][1], m22 = q[c2][1];
        return m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
      };
      double K[4] = {minor(1, 2, 3), -minor(0, 2, 3), minor(0, 1, 3), -minor(0, 1, 2)};
      auto cot = [&](int ap, int ia, int ib) {
        double ux = q[ia][0] - q[ap][0], uy = q[ia][1] - q[ap][1], vx = q[ib][0] - q[ap][0], vy = q[ib][1] - q[ap][1];
        double cr = ux * vy - uy * vx;
        return std::fabs(cr) > 1e-12 ? (ux * vx + uy * vy) / std::fabs(cr) : 0.0;
      };
      double want = cot(0, 1, 2) + cot(1, 0, 2);
      if (!(std::fabs(K[2]) > 1e-12)) continue;
      double a1 = std::fabs((q[1][0] - q[0][0]) * q[2][1]) / 2, a2 = std::fabs((q[1][0] - q[0][0]) * q[3][1]) / 2;
      Bend bd;
      bd.ids[0] = v0; bd.ids[1] = v1; bd.ids[2] = o2; bd.ids[3] = o3;
      for (int k = 0; k < 4; k++) bd.K[k] = (T)(float)(K[k] * (want / K[2]));
      bd.s = (T)(float)(3.0 * ((double)kv.second[0].second + (double)kv.second[1].second) / 2.0 / std::max(a1 + a2, 1e-9));
      bends.push_back(bd);
    }
    float meanEdge = edges.empty() ? 0.2f : edgeLenSum / (float)edges.size();
    elemMargin = std::min(0.01f, std::max(0.002f, 0.5f * maxPartR));
    float forceReach = 2.0f * maxPartR + elemMargin;
    datQueryRadius = std::max(1.5f * forceReach, 0.75f * meanEdge);
    float reach = (2.0f * maxPartR + elemMargin) / 0.7f;
    surfaceContactCellSize = std::max(0.02f, std::max(std::max(reach, meanEdge * 1.05f), maxElemR + maxPartR + elemMargin));
    elemCellSize = std::max(0.02f, std::max(std::max(reach, meanEdge * 1.05f), 2.0f * maxElemR + datQueryRadius));
    elemHashSize = nextPow2(std::max<uint32_t>(64u, 8u * nTris));
    uint64_t np = particles.size(), ne = edges.size();
    uint64_t cap = std::max<uint64_t>(1, 64 * np + 128 * ne);
    maxDatPairs = (uint32_t)std::min<uint64_t>(cap, 0xFFFFFFF0ull);
    maxSoft = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, 4ull * nTris + std::min<uint64_t>(cap, 8 * np + 8 * ne)),
                                           (1u << 24) - 1);
    if (isoOn) {
      // tracked-set capacity (GPUSolver.swift:439-440)
      maxSoft = (uint32_t)std::min<uint64_t>(
          std::max<uint64_t>(1, 4ull * particles.size() + 4ull * nTris + 4ull * numContactEdges), (1u << 24) - 1);
      isoTopology();
    }
  }

  // Static data of the isotropic mode (GPUSolver.swift:1336-1460): triangle
  // adjacency over shared edges (filled in ascending edge order), sorted
  // 2-rings, incident-edge lists in edge order, tracking state empty.
  void isoTopology() {
    std::map<uint64_t, std::pair<int64_t, int64_t>> edgeTris;
    for (uint32_t t = 0; t < nTris; t++) {
      const auto& id = tris[t];
      uint32_t ev[3][2] = {{id[0], id[1]}, {id[1], id[2]}, {id[0], id[2]}};
      for (auto& e : ev) {
        uint64_t k = (uint64_t)std::min(e[0], e[1]) << 32 | std::max(e[0], e[1]);
        auto it = edgeTris.find(k);
        if (it == edgeTris.end()) edgeTris[k] = {t, -1};
        else it->second.second = t;
      }
    }
    triAdj.assign(std::max<uint32_t>(1, nTris), {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu});
    std::vector<int> fill(std::max<uint32_t>(1, nTris), 0);
    for (const auto& kv : edgeTris) {
      if (kv.second.second < 0) continue;
      uint32_t a = (uint32_t)kv.second.first, b = (uint32_t)kv.second.second;
      if (fill[a] < 3) triAdj[a][fill[a]++] = b;
      if (fill[b] < 3) triAdj[b][fill[b]++] = a;
    }
    nbr2Start.assign(nb, 0);
    nbr2Count.assign(nb, 0);
    nbr2List.clear();
    for (uint32_t i = 0; i < nb; i++) {
      nbr2Start[i] = (uint32_t)nbr2List.size();
      if (nbrCount[i] == 0) continue;
      std::set<uint32_t> two;
      for (uint32_t k = 0; k < nbrCount[i]; k++) {
        uint32_t n1 = nbrList[nbrStart[i] + k];
        if (n1 != i) two.insert(n1);
        for (uint32_t q = 0; q < nbrCount[n1]; q++) {
          uint32_t n2 = nbrList[nbrStart[n1] + q];
          if (n2 != i) two.insert(n2);
        }
      }
      nbr2Count[i] = (uint32_t)two.size();
      nbr2List.insert(nbr2List.end(), two.begin(), two.end());
    }
    std::vector<std::vector<uint32_t>> ve(nb);
    for (uint32_t e = 0; e < edges.size(); e++) {
      ve[edges[e].first].push_back(e);
      ve[edges[e].second].push_back(e);
    }
    vEdgeStart.assign(nb, 0);
    vEdgeCount.assign(nb, 0);
    vEdgeList.clear();
    for (uint32_t i = 0; i < nb; i++) {
      vEdgeStart[i] = (uint32_t)vEdgeList.size();
      vEdgeCount[i] = (uint32_t)ve[i].size();
      vEdgeList.insert(vEdgeList.end(), ve[i].begin(), ve[i].end());
    }
    const std::array<uint32_t, 4> none{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    vtTrack.assign(nb, none);
    vtTrackPrev.assign(nb, none);
    eeTrack.assign(std::max<uint32_t>(1, numContactEdges), none);
    eeTrackPrev.assign(std::max<uint32_t>(1, numContactEdges), none);
    ogcBound.assign(nb, kOgcFar);
    ogcPrev.assign(nb, V<T>());
  }

  void rebuildBuckets() {
    buckets.assign(64, {});
    for (uint32_t v = 0; v < nb; v++) if (dyn(v)) buckets[colors[v] & 63u].push_back(v);
  }

  // -------------------------------------------------------------------------
  // broad phase (20_broadphase.metal:69-512) -- computed in float like the
  // kernels so cells and pair sets are comparable bit for bit
  // -------------------------------------------------------------------------
  V<T> center(uint32_t c) const { return pos[cOwner[c]] + qrot(rot[cOwner[c]], cLocal[c]); }
  static bool groupsOk(uint32_t a, uint32_t b) { return a == 0 || b == 0 || a == b; }
  static bool domainsOk(uint32_t ga, uint32_t gb, uint32_t sa, uint32_t sb) {
    if (!groupsOk(ga, gb)) return false;
    if (ga == 0 && gb != 0) return sb != 0;
    if (gb == 0 && ga != 0) return sa != 0;
    return true;
  }
  void cellOf(const V<T>& p, int& x, int& y, int& z) const {
    x = (int)std::floor(p.x / (T)cellSize);
    y = (int)std::floor(p.y / (T)cellSize);
    z = (int)std::floor(p.z / (T)cellSize);
  }
  bool excluded(uint32_t a, uint32_t b) const {
    return std::binary_search(excl.begin(), excl.end(), std::make_pair(a, b));
  }
  void broadphase() {
    cellCount.assign(hashSize, 0);
    cellStart.assign(hashSize, 0);
    std::vector<uint32_t> cellOfCol(hashed.size());
    std::vector<uint8_t> cellRigid(hashSize, 0);
    for (size_t k = 0; k < hashed.size(); k++) {
      uint32_t c = hashed[k];
      int x, y, z;
      cellOf(center(c), x, y, z);
      uint32_t h = cellHash32(x, y, z, cGroup[c], hashSize);
      cellOfCol[k] = h;
      cellCount[h]++;
      if (cRad[c] >= 0) cellRigid[h] = 1;
    }
    uint32_t run = 0;
    for (uint32_t h = 0; h < hashSize; h++) { cellStart[h] = run; run += cellCount[h]; }
    cellBodies.assign(hashed.size(), 0);
    std::vector<uint32_t> fill(hashSize, 0);
    std::vector<std::pair<uint32_t, uint32_t>> order;
    for (size_t k = 0; k < hashed.size(); k++) order.push_back({cellOfCol[k], hashed[k]});
    std::sort(order.begin(), order.end());  // each bucket sorted by collider id (bp_sort_cells)
    for (auto& o : order) cellBodies[cellStart[o.first] + fill[o.first]++] = o.second;
    // pairs, producer order = hashed then globals (bpProcessProducer)
    pairs.clear();
    pairCandidates = 0;
    auto accept = [&](uint32_t a, uint32_t b, bool canonical) {
      pairCandidates++;
      if (pairs.size() < maxPairs) pairs.push_back(canonical ? std::make_pair(std::min(a, b), std::max(a, b)) : std::make_pair(a, b));
    };
    auto test = [&](uint32_t a, uint32_t b, uint32_t ga, bool cellScan) -> bool {
      if (!domainsOk(cGroup[a], cGroup[b], cShared[a], cShared[b])) return false;
      uint32_t oa = cOwner[a], ob = cOwner[b];
      if (oa == ob) return false;
      if (!(mass[oa] > 0) && !(mass[ob] > 0)) return false;
      if (cellScan) { if (ga != 0 && clothGroup[ob] != 0) return false; }
      else { if (ga != 0 && clothGroup[ob] == ga) return false; }
      T ra = std::fabs(cRad[a]), rb = std::fabs(cRad[b]);
      V<T> dp = center(a) - center(b);
      T r = ra + rb;
      if (dotv(dp, dp) > r * r) return false;
      if (excluded(std::min(oa, ob), std::max(oa, ob))) return false;
      return true;
    };
    for (uint32_t a : hashed) {
      uint32_t ga = clothGroup[cOwner[a]];
      bool scan = !(ga != 0 && nHashedRigid == 0);
      if (scan) {
        int cx, cy, cz;
        cellOf(center(a), cx, cy, cz);
        std::vector<uint32_t> seen;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              uint32_t h = cellHash32(cx + dx, cy + dy, cz + dz, cGroup[a], hashSize);
              if (std::find(seen.begin(), seen.end(), h) != seen.end()) continue;
              seen.push_back(h);
              if (ga != 0 && !cellRigid[h]) continue;
              for (uint32_t k = cellStart[h]; k < cellStart[h] + cellCount[h]; k++) {
                uint32_t b = cellBodies[k];
                if (b <= a) continue;
                if (test(a, b, ga, true)) accept(a, b, false);
              }
            }
      }
      for (uint32_t b : globals)
        if (test(a, b, ga, false)) accept(a, b, true);
    }
    for (size_t gi = 0; gi < globals.size(); gi++) {
      uint32_t a = globals[gi];
      uint32_t ga = clothGroup[cOwner[a]];
      for (size_t gj = gi + 1; gj < globals.size(); gj++)
        if (test(a, globals[gj], ga, false)) accept(a, globals[gj], true);
    }
  }

  // -------------------------------------------------------------------------
  // narrow phase (30_narrowphase.metal:337-1322, sphere and box branches)
  // -------------------------------------------------------------------------
  struct Box { V<T> c, h, ax[3]; };
  struct Best { int type, ia, ib; T sep; V<T> n; bool valid; };
  struct Pt { V<T> xA, xB; uint32_t f; };

  T specCap(T ra, T rb) const {
    T r = std::min(std::fabs(ra), std::fabs(rb));
    return std::min(T(0.25), std::max(T(4) * (T)set.collision_margin, T(3) * r));
  }
  T detectMargin(const V<T>& n, const V<T>& rv, T ra, T rb) const {
    T ap = std::max(T(0), -dotv(rv, n));
    return (T)set.collision_margin + std::min(ap * (T)set.dt, specCap(ra, rb));
  }
  // kinematic spinner surface motion as a tangential target: the surface
  // displacement of each spinning body over the step, pre-divided by
  // (1 - alpha) because the solver discounts C0 by alpha
  // (30_narrowphase.metal:274-284); offA / offB are the contact points
  // relative to the body origins
  void spinShift(uint32_t ba, uint32_t bb, const V<T>& offA, const V<T>& offB, const V<T>& t1, const V<T>& t2,
                 V<T>& C0) const {
    if (spinBody.empty()) return;
    const T dt = (T)set.dt;
    V<T> sA = spinOn[ba] ? crossv(spinW[ba], offA) * dt : V<T>();
    V<T> sB = spinOn[bb] ? crossv(spinW[bb], offB) * dt : V<T>();
    V<T> rel = sB - sA;
    T k = T(1) / std::max(T(1) - (T)set.alpha, T(0.01));
    C0.y = C0.y - dotv(t1, rel) * k;
    C0.z = C0.z - dotv(t2, rel) * k;
  }
  V<T> penFloor(uint32_t a, uint32_t b) const {
    T big = std::numeric_limits<T>::max();
    T mA = mass[a], mB = mass[b];
    T mm = std::min(mA > 0 ? mA : big, mB > 0 ? mB : big);
    if (mm == big) return V<T>(1, 1, 1);
    T dt = set.dt;
    T k = std::min(T(1e10), std::max(T(1), mm / std::max(dt * dt, T(1e-12))));
    return V<T>(k, std::min(k, T(1e6)), std::min(k, T(1e6)));
  }
  static void basis(const V<T>& n, V<T>& t1, V<T>& t2) {
    t1 = std::fabs(n.x) > std::fabs(n.z) ? V<T>(-n.y, n.x, 0) : V<T>(0, -n.z, n.y);
    t1 = normv(t1);
    t2 = crossv(n, t1);
  }
  T combine(T a, T b) const {
    uint32_t m = set.friction_combine_mode;
    if (m == 1) return a * b;
    if (m == 2) return std::min(a, b);
    if (m == 3) return std::max(a, b);
    if (m == 4) return T(0.5) * (a + b);
    return std::sqrt(std::max(a * b, T(0)));
  }
  bool testAxis(const Box& A, const Box& B, const V<T>& delta, V<T> axis, int type, int ia, int ib, const V<T>& rv,
                T cap, Best& best) const {
    T l2 = dotv(axis, axis);
    if (l2 < T(1e-6)) return true;
    V<T> n = axis * (T(1) / std::sqrt(l2));
    if (dotv(n, delta) < 0) n = -n;
    T dist = std::fabs(dotv(delta, n));
    T rA = A.h.x * std::fabs(dotv(n, A.ax[0])) + A.h.y * std::fabs(dotv(n, A.ax[1])) + A.h.z * std::fabs(dotv(n, A.ax[2]));
    T rB = B.h.x * std::fabs(dotv(n, B.ax[0])) + B.h.y * std::fabs(dotv(n, B.ax[1])) + B.h.z * std::fabs(dotv(n, B.ax[2]));
    T sep = dist - (rA + rB);
    T allowed = std::min(std::max(T(0), -dotv(rv, n)) * (T)set.dt, cap);
    if (sep > allowed) return false;
    if (!best.valid || sep > best.sep) best = Best{type, ia, ib, sep, n, true};
    return true;
  }
  static V<T> support(const Box& b, const V<T>& d) {
    T sx = dotv(d, b.ax[0]) >= 0 ? T(1) : T(-1), sy = dotv(d, b.ax[1]) >= 0 ? T(1) : T(-1), sz = dotv(d, b.ax[2]) >= 0 ? T(1) : T(-1);
    return b.c + b.ax[0] * (b.h.x * sx) + b.ax[1] * (b.h.y * sy) + b.ax[2] * (b.h.z * sz);
  }
  static int clip(const V<T>* in, int cnt, const V<T>& n, T off, V<T>* out) {
    if (cnt <= 0) return 0;
    int oc = 0;
    V<T> a = in[cnt - 1];
    T da = dotv(n, a) - off;
    for (int i = 0; i < cnt; i++) {
      V<T> b = in[i];
      T db = dotv(n, b) - off;
      bool ai = da <= T(1e-5), bi = db <= T(1e-5);
      if (ai != bi) {
        T t = 0, den = da - db;
        if (std::fabs(den) > T(1e-6)) t = clampT(da / den, T(0), T(1));
        if (oc < 16) out[oc++] = a + (b - a) * t;
      }
      if (bi && oc < 16) out[oc++] = b;
      a = b;
      da = db;
    }
    return oc;
  }
  static bool addPt(std::vector<Pt>& cs, std::vector<V<T>>& mids, const V<T>& xA, const V<T>& xB, uint32_t f) {
    V<T> mid = (xA + xB) * T(0.5);
    for (auto& m : mids) { V<T> d = mid - m; if (dotv(d, d) < T(1e-6)) return false; }
    if (cs.size() >= 8) return false;
    cs.push_back(Pt{xA, xB, f});
    mids.push_back(mid);
    return true;
  }
  static void segSeg(V<T> p0, V<T> p1, V<T> q0, V<T> q1, V<T>& c0, V<T>& c1) {
    V<T> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    T a = dotv(d1, d1), e = dotv(d2, d2), f = dotv(d2, r), s = 0, t = 0;
    const T eps = T(1e-6);
    if (a <= eps && e <= eps) { c0 = p0; c1 = q0; return; }
    if (a <= eps) t = clampT(f / e, T(0), T(1));
    else {
      T c = dotv(d1, r);
      if (e <= eps) s = clampT(-c / a, T(0), T(1));
      else {
        T b = dotv(d1, d2), den = a * e - b * b;
        if (std::fabs(den) > eps) s = clampT((b * f - c * e) / den, T(0), T(1));
        t = (b * s + f) / e;
        if (t < 0) { t = 0; s = clampT(-c / a, T(0), T(1)); }
        else if (t > 1) { t = 1; s = clampT((b - c) / a, T(0), T(1)); }
      }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
  }
  static T half(const Box& b, int i) { return b.h[i]; }
  static void faceAxes(const Box& b, int k, V<T>& u, V<T>& v, T& eu, T& ev) {
    if (k == 0) { u = b.ax[1]; v = b.ax[2]; eu = b.h.y; ev = b.h.z; }
    else if (k == 1) { u = b.ax[0]; v = b.ax[2]; eu = b.h.x; ev = b.h.z; }
    else { u = b.ax[0]; v = b.ax[1]; eu = b.h.x; ev = b.h.y; }
  }
  static void supportEdge(const Box& b, int k, const V<T>& dir, V<T>& e0, V<T>& e1) {
    int a1 = (k + 1) % 3, a2 = (k + 2) % 3;
    T s1 = dotv(dir, b.ax[a1]) >= 0 ? T(1) : T(-1), s2 = dotv(dir, b.ax[a2]) >= 0 ? T(1) : T(-1);
    V<T> c = b.c + b.ax[a1] * (half(b, a1) * s1) + b.ax[a2] * (half(b, a2) * s2);
    e0 = c - b.ax[k] * half(b, k);
    e1 = c + b.ax[k] * half(b, k);
  }

  // ---- torus pairs (30_narrowphase.metal:292-336, 749-965) ----
  // The torus is the set of points within r of its spine, the circle of
  // radius R about the collider's local z axis through its center.
  struct Torus { V<T> c, axis, u, v; T R, r; };
  static Torus torusAt(const V<T>& p, const Q<T>& q, T R, T r) {
    Torus t;
    t.c = p;
    t.axis = qrot(q, V<T>(0, 0, 1));
    V<T> u = std::fabs(t.axis.x) > std::fabs(t.axis.z) ? V<T>(-t.axis.y, t.axis.x, 0) : V<T>(0, -t.axis.z, t.axis.y);
    t.u = normv(u);
    t.v = crossv(t.axis, t.u);
    t.R = R;
    t.r = r;
    return t;
  }
  // nearest spine point (the spine point along u for points on the axis)
  static V<T> spineNearest(const Torus& t, const V<T>& p) {
    V<T> d = p - t.c;
    d = d - t.axis * dotv(t.axis, d);
    T l = lenv(d);
    return l < T(1e-6) ? t.c + t.u * t.R : t.c + d / l * t.R;
  }
  // cos / sin of the 32 spine sample angles 2 pi i / 32, rounded to float
  static const float* torusTrig() {
    static float tab[64];
    static bool done = false;
    if (!done) {
      for (int i = 0; i < 32; i++) {
        const double a = 6.283185307179586476925 * (double)i / 32.0;
        tab[2 * i] = (float)std::cos(a);
        tab[2 * i + 1] = (float)std::sin(a);
      }
      done = true;
    }
    return tab;
  }
  void torusPair(uint32_t ba, uint32_t bb, uint32_t ia, uint32_t ib, bool tIsA, const V<T>& cA, const Q<T>& qA,
                 const V<T>& cB, const Q<T>& qB, T statF, T dynF, Manifold<T>& out) {
    struct Hit { V<T> xT, xO, n; T sep; uint32_t f; };  // n: other -> torus
    Hit hits[4];
    int nHits = 0;
    const uint32_t it = tIsA ? ia : ib, io = tIsA ? ib : ia;
    const Q<T> qO = tIsA ? qB : qA;
    const V<T> pT = tIsA ? cA : cB, pO = tIsA ? cB : cA;
    const Torus tor = torusAt(pT, tIsA ? qA : qB, cSize[it].x, cSize[it].y);
    const uint32_t other = cFlags[io] & 0xF;
    // detection reach: fixed, a fraction of the tube radius
    const T detectM = std::max((T)set.collision_margin, T(0.8) * tor.r);
    if (other == 1) {
      // torus - sphere: exact, through the nearest spine point
      const T rs = cSize[io].x * T(0.5);
      V<T> spine = spineNearest(tor, pO);
      V<T> d = pO - spine;
      T dist = lenv(d);
      if (dist <= tor.r + rs + detectM) {
        V<T> n = dist > T(1e-6) ? d / dist : V<T>(0, 0, 1);  // torus -> sphere
        hits[0] = Hit{spine + n * tor.r, pO - n * rs, -n, dist - tor.r - rs, 0u};
        nHits = 1;
      }
    } else {
      // box / capsule / torus: 32 spine samples, each with its exact nearest
      // point on the other shape; the deepest samples at least 4 apart are
      // refined by two alternating projections
      const bool isBox = other == 0, isCap = other == 3;
      const Q<T> qOc = qconj(qO);
      const V<T> halfO = cSize[io] * T(0.5);
      T otherR = 0;
      V<T> s0, s1;
      Torus t2{};
      if (isCap) {
        const T ch = cSize[io].x * T(0.5);
        V<T> ax = qrot(qO, V<T>(0, 0, 1));
        s0 = pO - ax * ch;
        s1 = pO + ax * ch;
        otherR = cSize[io].y;
      } else if (!isBox) {
        t2 = torusAt(pO, qO, cSize[io].x, cSize[io].y);
        otherR = t2.r;
      }
      auto nearestOther = [&](const V<T>& q) -> V<T> {
        if (isBox) {
          V<T> l = qrot(qOc, q - pO);
          V<T> cl(clampT(l.x, -halfO.x, halfO.x), clampT(l.y, -halfO.y, halfO.y), clampT(l.z, -halfO.z, halfO.z));
          return qrot(qO, cl) + pO;
        }
        if (isCap) {
          V<T> ab = s1 - s0;
          T t = clampT(dotv(q - s0, ab) / std::max(dotv(ab, ab), T(1e-12)), T(0), T(1));
          return s0 + ab * t;
        }
        return spineNearest(t2, q);
      };
      const float* trig = torusTrig();
      auto sample = [&](int i) { return tor.c + (tor.u * (T)trig[2 * i] + tor.v * (T)trig[2 * i + 1]) * tor.R; };
      T dists[32];
      bool used[32];
      for (int i = 0; i < 32; i++) {
        V<T> q = sample(i);
        dists[i] = lenv(q - nearestOther(q));
        used[i] = false;
      }
      for (int pick = 0; pick < 4; pick++) {
        int best = -1;
        T bd = std::numeric_limits<T>::max();
        for (int i = 0; i < 32; i++)
          if (!used[i] && dists[i] < bd) { bd = dists[i]; best = i; }
        if (best < 0 || bd - tor.r - otherR > detectM) break;
        for (int k = -3; k <= 3; k++) used[(best + k + 32) % 32] = true;
        V<T> q = sample(best);
        V<T> b = nearestOther(q);
        for (int r2 = 0; r2 < 2; r2++) {
          q = spineNearest(tor, b);
          b = nearestOther(q);
        }
        V<T> d = q - b;
        T dist = lenv(d);
        V<T> n, xO = b;
        if (dist < T(1e-6)) {
          if (isBox) {
            // spine point inside the box: out through the nearest face
            V<T> lb = qrot(qOc, q - pO);
            V<T> pen = halfO - V<T>(std::fabs(lb.x), std::fabs(lb.y), std::fabs(lb.z));
            int ax = 0;
            if (pen.y < pen[ax]) ax = 1;
            if (pen.z < pen[ax]) ax = 2;
            V<T> nl;
            nl[ax] = lb[ax] >= 0 ? T(1) : T(-1);
            n = qrot(qO, nl);
            xO = q + n * pen[ax];
          } else if (isCap) {
            n = normv(crossv(tor.axis, normv(s1 - s0)) + V<T>(T(1e-4), 0, 0));
          } else {
            n = normv(crossv(tor.axis, t2.axis) + V<T>(T(1e-4), 0, 0));
          }
        } else {
          if (dist - tor.r - otherR > detectM) continue;
          n = d / dist;  // other -> torus
        }
        hits[nHits++] = Hit{q - n * tor.r, isBox ? xO : xO + n * otherR, n, dist - tor.r - otherR, (uint32_t)best};
      }
    }
    if (nHits == 0) return;
    int deep = 0;
    T minSep = std::numeric_limits<T>::max();
    for (int h = 0; h < nHits; h++)
      if (hits[h].sep < minSep) { minSep = hits[h].sep; deep = h; }
    const V<T> nrm = tIsA ? hits[deep].n : -hits[deep].n;  // B -> A
    V<T> t1, t2;
    basis(nrm, t1, t2);
    auto pit = prevMap.find({ia, ib});
    const int prev = pit == prevMap.end() ? -1 : pit->second;
    const bool roundA = (cFlags[ia] & 0x20) != 0, roundB = (cFlags[ib] & 0x20) != 0;
    out.n = (uint32_t)nHits;
    out.flags = 1u | (roundA ? 2u : 0u) | (roundB ? 4u : 0u);
    out.nrm = nrm; out.t1 = t1; out.dynF = dynF; out.statF = statF;
    const T warm = (T)set.alpha * (T)set.gamma;
    const V<T> fl = penFloor(ba, bb);
    const T gam = set.gamma;
    for (int i = 0; i < nHits; i++) {
      const V<T> xAw = tIsA ? hits[i].xT : hits[i].xO, xBw = tIsA ? hits[i].xO : hits[i].xT;
      const V<T> rA_ = roundA ? xAw - pos[ba] : qrot(qconj(rot[ba]), xAw - pos[ba]);
      const V<T> rB_ = roundB ? xBw - pos[bb] : qrot(qconj(rot[bb]), xBw - pos[bb]);
      V<T> lam, pen;
      if (prev >= 0) {
        // nearest previous contact on A within 1.1 R carries its multipliers,
        // the tangential part turned into this frame's basis; never sticks
        const Manifold<T>& pm = prevManifolds[prev];
        T bestD = T(1.1) * cSize[it].x;
        int bj = -1;
        for (uint32_t j = 0; j < pm.n; j++) {
          T dj = lenv(pm.c[j].rA - rA_);
          if (dj < bestD) { bestD = dj; bj = (int)j; }
        }
        if (bj >= 0) {
          lam = pm.c[bj].lambda;
          pen = pm.c[bj].penalty;
          V<T> t2o = crossv(pm.nrm, pm.t1);
          V<T> lt = pm.t1 * lam.y + t2o * lam.z;
          lam.y = dotv(t1, lt);
          lam.z = dotv(t2, lt);
        }
      }
      const V<T> d = xAw - xBw;
      Contact<T>& c = out.c[i];
      c.rA = rA_; c.rB = rB_; c.feature = hits[i].f; c.stick = 0;
      c.C0 = V<T>(dotv(nrm, d) + (T)set.collision_margin, dotv(t1, d), dotv(t2, d));
      spinShift(ba, bb, xAw - pos[ba], xBw - pos[bb], t1, t2, c.C0);
      c.lambda = lam * warm;
      c.penalty = V<T>(clampT(pen.x * gam, fl.x, T(1e10)), clampT(pen.y * gam, fl.y, T(1e6)), clampT(pen.z * gam, fl.z, T(1e6)));
    }
  }

  void collide(uint32_t pi, Manifold<T>& out) {
    uint32_t ia = pairs[pi].first, ib = pairs[pi].second;
    uint32_t ba = cOwner[ia], bb = cOwner[ib];
    out = Manifold<T>{};
    out.a = ba; out.b = bb; out.ca = ia; out.cb = ib;
    V<T> cA = pos[ba] + qrot(rot[ba], cLocal[ia]), cB = pos[bb] + qrot(rot[bb], cLocal[ib]);
    Q<T> qA = qmul(rot[ba], cLocalRot[ia]), qB = qmul(rot[bb], cLocalRot[ib]);
    V<T> vA = vel[ba] + crossv(angVel[ba], cA - pos[ba]), vB = vel[bb] + crossv(angVel[bb], cB - pos[bb]);
    V<T> rv = vB - vA;
    Box A{cA, cSize[ia] * T(0.5), {qrot(qA, V<T>(1, 0, 0)), qrot(qA, V<T>(0, 1, 0)), qrot(qA, V<T>(0, 0, 1))}};
    Box B{cB, cSize[ib] * T(0.5), {qrot(qB, V<T>(1, 0, 0)), qrot(qB, V<T>(0, 1, 0)), qrot(qB, V<T>(0, 0, 1))}};
    V<T> delta = B.c - A.c;
    bool sphA = (cFlags[ia] & 0xF) == 1, sphB = (cFlags[ib] & 0xF) == 1;
    T statF = combine(cStatF[ia], cStatF[ib]), dynF = combine(cDynF[ia], cDynF[ib]);
    T warm = (T)set.alpha * (T)set.gamma;
    T gam = set.gamma;
    std::vector<Pt> cs;
    std::vector<V<T>> mids;
    V<T> nrm;
    auto prevIndex = [&]() -> int {
      auto it = prevMap.find({ia, ib});
      return it == prevMap.end() ? -1 : it->second;
    };
    V<T> fl = penFloor(ba, bb);
    auto warmPen = [&](V<T> p) {
      return V<T>(clampT(p.x * gam, fl.x, T(1e10)), clampT(p.y * gam, fl.y, T(1e6)), clampT(p.z * gam, fl.z, T(1e6)));
    };
    bool capA = (cFlags[ia] & 0xF) == 3, capB = (cFlags[ib] & 0xF) == 3;
    bool hullA = (cFlags[ia] & 0xF) == 4, hullB = (cFlags[ib] & 0xF) == 4;
    // hulls collide with boxes only (30_narrowphase.metal:404-410)
    if ((hullA || hullB) && ((hullA && hullB) || (cFlags[hullA ? ib : ia] & 0xF) != 0)) return;
    bool torA = (cFlags[ia] & 0xF) == 2, torB = (cFlags[ib] & 0xF) == 2;
    if (torA || torB) {
      torusPair(ba, bb, ia, ib, torA, A.c, qA, B.c, qB, statF, dynF, out);
      return;
    }
    if (capA || capB || sphA || sphB) {
      T rA = cSize[ia].x * T(0.5), rB = cSize[ib].x * T(0.5);
      T bestDRound;
      if (capA || capB) {
        // capsule pairs (30_narrowphase.metal:510-745)
        bool cIsA = capA;
        uint32_t ic = cIsA ? ia : ib, io = cIsA ? ib : ia;
        Q<T> qC = cIsA ? qA : qB, qO = cIsA ? qB : qA;
        V<T> pC = cIsA ? A.c : B.c, pO = cIsA ? B.c : A.c;
        T rc = cSize[ic].y;
        V<T> axC = qrot(qC, V<T>(0, 0, 1));
        V<T> c0 = pC - axC * (cSize[ic].x * T(0.5)), c1 = pC + axC * (cSize[ic].x * T(0.5));
        uint32_t other = cFlags[io] & 0xF;
        V<T> relCO = cIsA ? (vB - vA) : (vA - vB);
        V<T> nCO(0, 0, 1);
        std::vector<Pt> hits;  // xA on the capsule, xB on the other shape
        auto onSeg = [](const V<T>& p, const V<T>& a, const V<T>& b) {
          V<T> ab = b - a;
          T t = clampT(dotv(p - a, ab) / std::max(dotv(ab, ab), T(1e-12)), T(0), T(1));
          return a + ab * t;
        };
        if (other == 1) {
          T rs = cSize[io].x * T(0.5);
          V<T> q = onSeg(pO, c0, c1);
          V<T> d = pO - q;
          T dist = lenv(d);
          V<T> nT = dist > T(1e-6) ? d / dist : V<T>(0, 0, 1);
          if (dist <= rc + rs + detectMargin(nT, relCO, cRad[ic], cRad[io])) {
            nCO = nT;
            hits.push_back(Pt{q + nCO * rc, pO - nCO * rs, 0});
          }
        } else if (other == 3) {
          T ro = cSize[io].y;
          V<T> axO = qrot(qO, V<T>(0, 0, 1));
          V<T> o0 = pO - axO * (cSize[io].x * T(0.5)), o1 = pO + axO * (cSize[io].x * T(0.5));
          V<T> qa, qb;
          segSeg(c0, c1, o0, o1, qa, qb);
          V<T> d = qb - qa;
          T dist = lenv(d);
          V<T> nT = dist > T(1e-6) ? d / dist : V<T>(0, 0, 1);
          if (dist <= rc + ro + detectMargin(nT, relCO, cRad[ic], cRad[io])) {
            nCO = nT;
            hits.push_back(Pt{qa + nCO * rc, qb - nCO * ro, 0});
          }
        } else {
          Q<T> qOc = qconj(qO);
          V<T> h = cSize[io] * T(0.5);
          V<T> nAcc;
          for (int seed = 0; seed < 3 && hits.size() < 3; seed++) {
            V<T> q = c0 + (c1 - c0) * (T(seed) * T(0.5));
            V<T> bw = q;
            for (int k = 0; k < 5; k++) {
              V<T> l = qrot(qOc, q - pO);
              V<T> lb(clampT(l.x, -h.x, h.x), clampT(l.y, -h.y, h.y), clampT(l.z, -h.z, h.z));
              bw = qrot(qO, lb) + pO;
              q = onSeg(bw, c0, c1);
            }
            V<T> d = bw - q;
            T dist = lenv(d);
            V<T> n;
            if (dist < T(1e-6)) {
              V<T> lb = qrot(qOc, q - pO);
              V<T> pen = h - V<T>(std::fabs(lb.x), std::fabs(lb.y), std::fabs(lb.z));
              int ax = 0;
              if (pen.y < pen[ax]) ax = 1;
              if (pen.z < pen[ax]) ax = 2;
              V<T> nl;
              nl[ax] = lb[ax] >= 0 ? T(1) : T(-1);
              n = -qrot(qO, nl);
              bw = q - n * pen[ax];
            } else {
              n = d / dist;
              if (dist - rc > detectMargin(n, relCO, cRad[ic], cRad[io])) continue;
            }
            bool dup = false;
            for (const Pt& hp : hits)
              if (lenv(hp.xA - (q + n * rc)) < T(0.4) * rc + T(0.05)) { dup = true; break; }
            if (dup) continue;
            hits.push_back(Pt{q + n * rc, bw, (uint32_t)seed});
            nAcc = nAcc + n;
          }
          if (!hits.empty()) nCO = normv(nAcc);
        }
        if (hits.empty()) return;
        nrm = cIsA ? -nCO : nCO;
        for (const Pt& hp : hits) cs.push_back(Pt{cIsA ? hp.xA : hp.xB, cIsA ? hp.xB : hp.xA, hp.f});
        bestDRound = T(0.6) * (rc + T(0.2));
      } else if (sphA && sphB) {
        V<T> d = A.c - B.c;
        T dist = lenv(d);
        V<T> nT = dist > T(1e-9) ? d / dist : V<T>(0, 0, 1);
        if (dist > rA + rB + detectMargin(nT, vA - vB, rA, rB)) return;
        nrm = nT;
        cs.push_back(Pt{A.c - nrm * rA, B.c + nrm * rB, 0});
      } else {
        bool sIsA = sphA;
        T r = sIsA ? rA : rB;
        const Box& box = sIsA ? B : A;
        Q<T> qBox = sIsA ? qB : qA;
        V<T> sc = sIsA ? A.c : B.c;
        V<T> local = qrot(qconj(qBox), sc - box.c);
        V<T> cl(clampT(local.x, -box.h.x, box.h.x), clampT(local.y, -box.h.y, box.h.y), clampT(local.z, -box.h.z, box.h.z));
        V<T> nL, qL;
        if (cl.x == local.x && cl.y == local.y && cl.z == local.z) {
          V<T> pen = box.h - V<T>(std::fabs(local.x), std::fabs(local.y), std::fabs(local.z));
          int ax = 0;
          if (pen.y < pen[ax]) ax = 1;
          if (pen.z < pen[ax]) ax = 2;
          nL = V<T>();
          nL[ax] = local[ax] >= 0 ? T(1) : T(-1);
          qL = local;
          qL[ax] = nL[ax] * box.h[ax];
        } else {
          V<T> d = local - cl;
          T dist = lenv(d);
          V<T> nLT = d / std::max(dist, T(1e-9));
          V<T> nWT = qrot(qBox, nLT);
          V<T> vS = sIsA ? vA : vB, vX = sIsA ? vB : vA;
          if (dist > r + detectMargin(nWT, vS - vX, r, sIsA ? cRad[ib] : cRad[ia])) return;
          nL = nLT;
          qL = cl;
        }
        V<T> nW = qrot(qBox, nL), qW = qrot(qBox, qL) + box.c, xS = sc - nW * r;
        nrm = sIsA ? nW : -nW;
        cs.push_back(Pt{sIsA ? xS : qW, sIsA ? qW : xS, 0});
      }
      if (!(capA || capB)) bestDRound = T(0.5) * std::max(rA, rB);
      V<T> t1, t2;
      basis(nrm, t1, t2);
      int prev = prevIndex();
      bool roundA = (cFlags[ia] & 0x20) != 0, roundB = (cFlags[ib] & 0x20) != 0;
      out.n = (uint32_t)cs.size();
      out.flags = 1u | (roundA ? 2u : 0u) | (roundB ? 4u : 0u);
      out.nrm = nrm; out.t1 = t1; out.dynF = dynF; out.statF = statF;
      for (size_t i = 0; i < cs.size(); i++) {
        V<T> rA_ = roundA ? cs[i].xA - pos[ba] : qrot(qconj(rot[ba]), cs[i].xA - pos[ba]);
        V<T> rB_ = roundB ? cs[i].xB - pos[bb] : qrot(qconj(rot[bb]), cs[i].xB - pos[bb]);
        V<T> lam, pen;
        T stick = 0;
        if (prev >= 0) {
          const Manifold<T>& pm = prevManifolds[prev];
          T bestD = bestDRound;
          int bj = -1;
          for (uint32_t j = 0; j < pm.n; j++) {
            T dj = lenv(pm.c[j].rA - rA_);
            if (dj < bestD) { bestD = dj; bj = (int)j; }
          }
          if (bj >= 0) {
            lam = pm.c[bj].lambda;
            pen = pm.c[bj].penalty;
            V<T> t2o = crossv(pm.nrm, pm.t1);
            V<T> lt = pm.t1 * lam.y + t2o * lam.z;
            lam.y = dotv(t1, lt);
            lam.z = dotv(t2, lt);
            bool pA = cRad[ia] < 0, pB = cRad[ib] < 0;
            bool keep = (capA || capB) ? (!roundA && !roundB) : ((pA || !roundA) && (pB || !roundB));
            if (keep) {
              stick = pm.c[bj].stick;
              if (stick != 0) { rA_ = pm.c[bj].rA; rB_ = pm.c[bj].rB; }
            }
          }
        }
        V<T> xAw = roundA ? pos[ba] + rA_ : xformv(pos[ba], rot[ba], rA_);
        V<T> xBw = roundB ? pos[bb] + rB_ : xformv(pos[bb], rot[bb], rB_);
        V<T> d = xAw - xBw;
        Contact<T>& c = out.c[i];
        c.rA = rA_; c.rB = rB_; c.feature = cs[i].f; c.stick = stick;
        c.C0 = V<T>(dotv(nrm, d) + (T)set.collision_margin, dotv(t1, d), dotv(t2, d));
        spinShift(ba, bb, xAw - pos[ba], xBw - pos[bb], t1, t2, c.C0);
        c.lambda = lam * warm;
        c.penalty = warmPen(pen);
      }
      return;
    }
    if (hullA || hullB) {
      // convex hull vs box (30_narrowphase.metal:189-247, 411-497)
      bool hIsA = hullA;
      uint32_t ih = hIsA ? ia : ib;
      Q<T> qH = hIsA ? qA : qB, qBox = hIsA ? qB : qA;
      V<T> pH = hIsA ? A.c : B.c;
      const Box& box = hIsA ? B : A;
      V<T> rel = hIsA ? (vA - vB) : (vB - vA);
      T hullR = std::fabs(cRad[ih]), boxR = std::fabs(cRad[hIsA ? ib : ia]);
      struct Cand { V<T> xHull, xBox, n; T sep; uint32_t f; };
      auto vertex = [&](uint32_t vid, Cand& c) -> bool {
        V<T> xw = pH + qrot(qH, cHull[ih][vid]);
        V<T> l = qrot(qconj(qBox), xw - box.c);
        V<T> cl(clampT(l.x, -box.h.x, box.h.x), clampT(l.y, -box.h.y, box.h.y), clampT(l.z, -box.h.z, box.h.z));
        V<T> nL, bL;
        T sep;
        bool inside = l.x >= -box.h.x && l.y >= -box.h.y && l.z >= -box.h.z && l.x <= box.h.x && l.y <= box.h.y &&
                      l.z <= box.h.z;
        if (inside) {
          V<T> clr = box.h - V<T>(std::fabs(l.x), std::fabs(l.y), std::fabs(l.z));
          int ax = 0;
          if (clr.y < clr[ax]) ax = 1;
          if (clr.z < clr[ax]) ax = 2;
          nL = V<T>();
          nL[ax] = l[ax] >= 0 ? T(1) : T(-1);
          bL = l;
          bL[ax] = nL[ax] * box.h[ax];
          sep = -clr[ax];
        } else {
          V<T> dd = l - cl;
          T dist = lenv(dd);
          if (dist < T(1e-8)) return false;
          nL = dd / dist;
          if (dist > detectMargin(qrot(qBox, nL), rel, hullR, boxR)) return false;
          bL = cl;
          sep = dist;
        }
        c = Cand{xw, qrot(qBox, bL) + box.c, qrot(qBox, nL), sep, (4u << 24) | (vid & 0x00FFFFFFu)};
        return true;
      };
      const uint32_t nv = (uint32_t)cHull[ih].size();
      Cand best{};
      bool have = false;
      for (uint32_t v = 0; v < nv; v++) {
        Cand c;
        if (!vertex(v, c)) continue;
        if (!have || c.sep < best.sep - T(1e-8) || (std::fabs(c.sep - best.sep) <= T(1e-8) && c.f < best.f)) {
          best = c;
          have = true;
        }
      }
      if (!have) return;
      std::vector<Cand> chosen{best};
      T band = T(0.75) * (T)set.collision_margin;
      while (chosen.size() < 8) {
        Cand next{};
        bool haveNext = false;
        T bestSpread = T(-1);
        for (uint32_t v = 0; v < nv; v++) {
          Cand c;
          if (!vertex(v, c)) continue;
          if (c.sep > best.sep + band || dotv(c.n, best.n) < T(0.98)) continue;
          bool dup = false;
          T minSpread = std::numeric_limits<T>::max();
          for (const Cand& ch : chosen) {
            if (c.f == ch.f) { dup = true; break; }
            V<T> dd = c.xHull - ch.xHull;
            dd = dd - best.n * dotv(best.n, dd);
            minSpread = std::min(minSpread, dotv(dd, dd));
          }
          if (dup) continue;
          if (!haveNext || minSpread > bestSpread + T(1e-10) ||
              (std::fabs(minSpread - bestSpread) <= T(1e-10) && c.f < next.f)) {
            next = c;
            bestSpread = minSpread;
            haveNext = true;
          }
        }
        if (!haveNext || bestSpread < T(1e-6)) break;
        chosen.push_back(next);
      }
      for (const Cand& ch : chosen) cs.push_back(Pt{hIsA ? ch.xHull : ch.xBox, hIsA ? ch.xBox : ch.xHull, ch.f});
      nrm = hIsA ? best.n : -best.n;
    } else {
    // box-box SAT
    Best bf{0, 0, 0, -std::numeric_limits<T>::max(), V<T>(), false}, be = bf;
    T cap = specCap(cRad[ia], cRad[ib]);
    bool sepd = false;
    for (int i = 0; i < 3 && !sepd; i++) if (!testAxis(A, B, delta, A.ax[i], 0, i, -1, rv, cap, bf)) sepd = true;
    for (int i = 0; i < 3 && !sepd; i++) if (!testAxis(A, B, delta, B.ax[i], 1, -1, i, rv, cap, bf)) sepd = true;
    for (int i = 0; i < 3 && !sepd; i++)
      for (int j = 0; j < 3 && !sepd; j++)
        if (!testAxis(A, B, delta, crossv(A.ax[i], B.ax[j]), 2, i, j, rv, cap, be)) sepd = true;
    if (sepd || !bf.valid) return;
    Best best = bf;
    if (be.valid && T(0.95) * be.sep > bf.sep + T(0.01)) best = be;
    if (best.type == 2) {
      V<T> a0, a1, b0, b1, xA, xB;
      supportEdge(A, best.ia, best.n, a0, a1);
      supportEdge(B, best.ib, -best.n, b0, b1);
      segSeg(a0, a1, b0, b1, xA, xB);
      uint32_t f = (2u << 24) | ((uint32_t)(best.ia & 0xFF) << 8) | (uint32_t)(best.ib & 0xFF);
      addPt(cs, mids, xA, xB, f);
      if (cs.empty()) addPt(cs, mids, support(A, best.n), support(B, -best.n), f);
    } else {
      bool refIsA = best.type == 0;
      int ra = refIsA ? best.ia : best.ib;
      const Box& ref = refIsA ? A : B;
      const Box& inc = refIsA ? B : A;
      V<T> ro = refIsA ? best.n : -best.n;
      T sR = dotv(ro, ref.ax[ra]) >= 0 ? T(1) : T(-1);
      V<T> rN = ref.ax[ra] * sR, rC = ref.c + rN * half(ref, ra);
      V<T> rU, rV;
      T eU, eV;
      faceAxes(ref, ra, rU, rV, eU, eV);
      int ia2 = 0;
      T bd = -std::numeric_limits<T>::max();
      for (int i = 0; i < 3; i++) { T dd = std::fabs(dotv(inc.ax[i], rN)); if (dd > bd) { bd = dd; ia2 = i; } }
      T sI = dotv(inc.ax[ia2], rN) > 0 ? T(-1) : T(1);
      V<T> iN = inc.ax[ia2] * sI, iC = inc.c + iN * half(inc, ia2);
      V<T> iU, iV;
      T fU, fV;
      faceAxes(inc, ia2, iU, iV, fU, fV);
      V<T> p0[16], p1[16];
      p0[0] = iC + iU * fU + iV * fV;
      p0[1] = iC - iU * fU + iV * fV;
      p0[2] = iC - iU * fU - iV * fV;
      p0[3] = iC + iU * fU - iV * fV;
      int n = 4;
      n = clip(p0, n, rU, dotv(rU, rC) + eU, p1);
      n = clip(p1, n, -rU, dotv(-rU, rC) + eU, p0);
      n = clip(p0, n, rV, dotv(rV, rC) + eV, p1);
      n = clip(p1, n, -rV, dotv(-rV, rC) + eV, p0);
      uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | ((uint32_t)(ra & 0xFF) << 16) | ((uint32_t)(ia2 & 0xFF) << 8);
      T slop = T(1e-5) + std::max(T(0), best.sep);
      for (int i = 0; i < n && cs.size() < 8; i++) {
        V<T> pI = p0[i];
        T dd = dotv(pI - rC, rN);
        if (dd > slop) continue;
        V<T> pR = pI - rN * dd;
        addPt(cs, mids, refIsA ? pR : pI, refIsA ? pI : pR, prefix | (uint32_t)(i & 0xFF));
      }
      if (cs.empty()) addPt(cs, mids, support(A, best.n), support(B, -best.n), prefix);
    }
    nrm = -best.n;
    }  // box-box SAT
    V<T> t1, t2;
    basis(nrm, t1, t2);
    int prev = prevIndex();
    out.n = (uint32_t)cs.size();
    out.flags = 1u;
    out.nrm = nrm; out.t1 = t1; out.dynF = dynF; out.statF = statF;
    for (size_t i = 0; i < cs.size(); i++) {
      V<T> rA = qrot(qconj(rot[ba]), cs[i].xA - pos[ba]);
      V<T> rB = qrot(qconj(rot[bb]), cs[i].xB - pos[bb]);
      V<T> lam, pen;
      T stick = 0;
      if (prev >= 0) {
        const Manifold<T>& pm = prevManifolds[prev];
        for (uint32_t j = 0; j < pm.n; j++) {
          if (pm.c[j].feature == cs[i].f) {
            lam = pm.c[j].lambda; pen = pm.c[j].penalty; stick = pm.c[j].stick;
            if (stick != 0) { rA = pm.c[j].rA; rB = pm.c[j].rB; }
            break;
          }
        }
      }
      V<T> xAw = xformv(pos[ba], rot[ba], rA), xBw = xformv(pos[bb], rot[bb], rB);
      V<T> d = xAw - xBw;
      Contact<T>& c = out.c[i];
      c.rA = rA; c.rB = rB; c.feature = cs[i].f; c.stick = stick;
      c.C0 = V<T>(dotv(nrm, d) + (T)set.collision_margin, dotv(t1, d), dotv(t2, d));
      spinShift(ba, bb, xAw - pos[ba], xBw - pos[bb], t1, t2, c.C0);
      c.lambda = lam * warm;
      c.penalty = warmPen(pen);
    }
  }

  // -------------------------------------------------------------------------
  // warm start (40_solver.metal:546-722)
  // -------------------------------------------------------------------------
  static T twistOf(const Q<T>& qAr, const Q<T>& qB, const V<T>& axis) {
    Q<T> r = qmul(qinv(qAr), qB);
    if (r.w < 0) r = Q<T>{-r.x, -r.y, -r.z, -r.w};
    return T(2) * std::atan2(dotv(V<T>(r.x, r.y, r.z), axis), r.w);
  }
  static T wrapPi(T a) { return a - T(6.2831853) * std::floor((a + T(3.14159265)) / T(6.2831853)); }

  void warmstartConstraints() {
    T warm = (T)set.alpha * (T)set.gamma, gam = set.gamma;
    for (auto& sp : springs) {
      if (!sp.hard) continue;
      V<T> pA = xformv(pos[sp.a], rot[sp.a], sp.rA), pB = xformv(pos[sp.b], rot[sp.b], sp.rB);
      sp.C0 = T(0.5) * std::max(lenv(pA - pB) - sp.rest, T(0));
      sp.lambda *= warm;
      sp.penalty = std::min(clampT(sp.penalty * gam, T(1), T(1e10)), sp.stiffness);
    }
    for (auto& j : joints) {
      if (j.broken) continue;
      V<T> pA = j.a == WORLD ? j.rA : xformv(pos[j.a], rot[j.a], j.rA);
      V<T> pB = xformv(pos[j.b], rot[j.b], j.rB);
      Q<T> qA = j.a == WORLD ? Q<T>{0, 0, 0, 1} : rot[j.a];
      j.C0Lin = pA - pB;
      if (j.hinge) {
        V<T> aB = qrot(rot[j.b], j.axis), aA = qrot(qmul(qA, j.restRel), j.axis);
        j.C0Ang = crossv(aA, aB) * j.torqueArm;
        if (j.armature > 0) {
          T tw = twistOf(qmul(qA, j.restRel), rot[j.b], j.axis);
          V<T> wA = j.a == WORLD ? V<T>() : angVel[j.a];
          j.predTwist = tw + (T)set.dt * dotv(angVel[j.b] - wA, aB);
        }
      } else {
        j.C0Ang = qsub(qmul(qA, j.restRel), rot[j.b]) * j.torqueArm;
      }
      if (j.motorTorque > 0) {
        j.motorZ = 0;
        j.effort = 0;
        uint32_t mode = j.flags & (3u << 4);
        if (mode == (2u << 4) || mode == (3u << 4)) {
          Q<T> qAr = j.a == WORLD ? j.restRel : qmul(rot[j.a], j.restRel);
          T tw = twistOf(qAr, rot[j.b], j.axis);
          V<T> axW = qrot(rot[j.b], j.axis);
          V<T> wA = j.a == WORLD ? V<T>() : angVel[j.a];
          T rate = dotv(angVel[j.b] - wA, axW);
          if (mode == (3u << 4)) j.effort = clampT(j.kd * (j.motorTarget - rate), -j.motorTorque, j.motorTorque);
          else j.effort = clampT(j.kp * (j.motorTarget - tw) - j.kd * rate, -j.motorTorque, j.motorTorque);
        }
      }
      j.lambdaLin = j.lambdaLin * warm;
      j.lambdaAng = j.lambdaAng * warm;
      for (int k = 0; k < 3; k++) {
        j.penaltyLin[k] = std::min(clampT(j.penaltyLin[k] * gam, T(1), T(1e10)), j.stiffLin);
        j.penaltyAng[k] = std::min(clampT(j.penaltyAng[k] * gam, T(1), T(1e10)), j.stiffAng);
      }
    }
  }

  void warmstartBodies() {
    T dt = set.dt, g0 = set.gravity;
    for (uint32_t i = 0; i < nb; i++) {
      T m = mass[i];
      T g = g0 * gravityScale[i];
      V<T> inert = pos[i] + vel[i] * dt;
      if (m > 0) inert = inert + V<T>(0, 0, g) * (dt * dt);
      inertLin[i] = inert;
      inertAng[i] = qaddw(rot[i], angVel[i] * dt);
      V<T> acc = (vel[i] - prevVel[i]) / dt;
      T gs = g0 < 0 ? T(-1) : (g0 > 0 ? T(1) : T(0));
      T w = clampT(acc.z * gs / std::fabs(g0), T(0), T(1));
      if (!std::isfinite(w)) w = 0;
      initLin[i] = pos[i];
      initAng[i] = rot[i];
      lo[i] = V<T>();
      const bool isoPart = isoOn && radiusSigned[i] < 0;
      if (isoPart) ogcPrev[i] = pos[i];
      if (m > 0) {
        V<T> guess = pos[i] + vel[i] * dt + V<T>(0, 0, g) * (w * dt * dt);
        // the predictor stays within the clearance bound of its particle
        // (40_solver.metal:706-716)
        if (isoPart && ogcBound[i] < T(1e37)) {
          T bv = std::max(T(0.45) * std::sqrt(ogcBound[i]), T(0.2) * std::fabs(radiusSigned[i]));
          V<T> tot = guess - pos[i];
          T t2 = dotv(tot, tot);
          if (t2 > bv * bv) guess = pos[i] + tot * (bv / std::sqrt(t2));
        }
        pos[i] = guess;
        rot[i] = inertAng[i];
      }
    }
    snapLin = pos;
    snapAng = rot;
  }

  // -------------------------------------------------------------------------
  // per-body system (40_solver.metal:724-1360, 1384-1495)
  // -------------------------------------------------------------------------
  struct Sys {
    M<T> ll, aa, cr;
    V<T> gl, ga;
  };
  V<T> pL(uint32_t b, uint32_t self) const {
    return (b != self && b != WORLD && colors[b] == colors[self] && dyn(b)) ? snapLin[b] : pos[b];
  }
  Q<T> pA(uint32_t b, uint32_t self) const {
    return (b != self && b != WORLD && colors[b] == colors[self] && dyn(b)) ? snapAng[b] : rot[b];
  }

  void stampTet(const Tet<T>& t, uint32_t self, Sys& s) const {
    // edges from vertex 0, high and low parts differenced separately
    auto edge = [&](uint32_t k) { return (pos[t.ids[k]] - pos[t.ids[0]]) + (lo[t.ids[k]] - lo[t.ids[0]]); };
    T muV = t.muV, lamV = t.lamV;
    T aT = T(1) + muV / std::max(lamV, T(1e-9));
    V<T> d0 = edge(1), d1 = edge(2), d2 = edge(3);
    V<T> f[3];
    for (int c = 0; c < 3; c++) f[c] = d0 * t.r[0][c] + d1 * t.r[1][c] + d2 * t.r[2][c];
    V<T> j0 = crossv(f[1], f[2]), j1 = crossv(f[2], f[0]), j2 = crossv(f[0], f[1]);
    T J = dotv(f[0], j0);
    T s2 = lamV * (J - aT);
    V<T> p0 = f[0] * muV + j0 * s2, p1 = f[1] * muV + j1 * s2, p2 = f[2] * muV + j2 * s2;
    V<T> w;
    if (self == t.ids[1]) w = t.r[0];
    else if (self == t.ids[2]) w = t.r[1];
    else if (self == t.ids[3]) w = t.r[2];
    else w = -(t.r[0] + t.r[1] + t.r[2]);
    s.gl = s.gl + p0 * w.x + p1 * w.y + p2 * w.z;
    V<T> g = j0 * w.x + j1 * w.y + j2 * w.z;
    s.ll = s.ll + M<T>::diag(V<T>(1, 1, 1) * (muV * dotv(w, w))) + M<T>::outer(g, g) * lamV;
  }

  V<T> contactC(const Manifold<T>& m, int i, const V<T>& dqAL, const V<T>& dqAA, const V<T>& dqBL, const V<T>& dqBA,
                const V<T>& rAW, const V<T>& rBW, T& fs, T& bnd, V<T>& F) const {
    V<T> n = m.nrm, t1 = m.t1, t2 = crossv(n, t1);
    const Contact<T>& c = m.c[i];
    V<T> C = c.C0 * (T(1) - (T)set.alpha);
    C = C + V<T>(dotv(n, dqAL), dotv(t1, dqAL), dotv(t2, dqAL)) - V<T>(dotv(n, dqBL), dotv(t1, dqBL), dotv(t2, dqBL));
    C = C + V<T>(dotv(crossv(rAW, n), dqAA), dotv(crossv(rAW, t1), dqAA), dotv(crossv(rAW, t2), dqAA));
    C = C - V<T>(dotv(crossv(rBW, n), dqBA), dotv(crossv(rBW, t1), dqBA), dotv(crossv(rBW, t2), dqBA));
    F = mulv(c.penalty, C) + c.lambda;
    F.x = std::min(F.x, T(0));
    bnd = std::fabs(F.x) * m.statF;
    fs = std::sqrt(F.y * F.y + F.z * F.z);
    if (fs > bnd && fs > 0) {
      T dynB = std::fabs(F.x) * m.dynF;
      F.y = F.y * (dynB / fs);
      F.z = F.z * (dynB / fs);
    }
    return C;
  }

  void stampManifold(const Manifold<T>& m, uint32_t self, Sys& s) const {
    bool isA = self == m.a;
    bool rAr = (m.flags & 2u) != 0, rBr = (m.flags & 4u) != 0;
    V<T> pAL = pL(m.a, self), pBL = pL(m.b, self);
    Q<T> qA = pA(m.a, self), qB = pA(m.b, self);
    V<T> dqAL = pAL - initLin[m.a], dqBL = pBL - initLin[m.b];
    if (isA) dqAL = dqAL + lo[self];   // the stamping body's own low part
    else dqBL = dqBL + lo[self];
    V<T> dqAA = qsub(qA, initAng[m.a]), dqBA = qsub(qB, initAng[m.b]);
    V<T> n = m.nrm, t1 = m.t1, t2 = crossv(n, t1);
    T sg = isA ? T(1) : T(-1);
    for (uint32_t i = 0; i < m.n; i++) {
      V<T> rAW = rAr ? m.c[i].rA : qrot(qA, m.c[i].rA);
      V<T> rBW = rBr ? m.c[i].rB : qrot(qB, m.c[i].rB);
      T fs, bnd;
      V<T> F;
      contactC(m, (int)i, dqAL, dqAA, dqBL, dqBA, rAW, rBW, fs, bnd, F);
      V<T> jl[3] = {n * sg, t1 * sg, t2 * sg};
      V<T> rW = isA ? rAW : rBW;
      V<T> K = m.c[i].penalty;
      for (int k = 0; k < 3; k++) {
        V<T> ja = crossv(rW, jl[k]);
        s.ll = s.ll + M<T>::outer(jl[k], jl[k]) * K[k];
        s.aa = s.aa + M<T>::outer(ja, ja) * K[k];
        s.cr = s.cr + M<T>::outer(ja, jl[k]) * K[k];
        s.gl = s.gl + jl[k] * F[k];
        s.ga = s.ga + ja * F[k];
      }
    }
  }

  static M<T> geomBS(int k, const V<T>& v) {
    M<T> m = M<T>::diag(V<T>(-v[k], -v[k], -v[k]));
    for (int r = 0; r < 3; r++) m.a[r][k] += v[r];
    return m;
  }
  void stampJoint(const Joint<T>& j, uint32_t self, Sys& s) const {
    bool isA = self == j.a;
    T alpha = set.alpha, dt = set.dt;
    if (dotv(j.penaltyLin, j.penaltyLin) > 0) {
      V<T> pAw = j.a == WORLD ? j.rA : xformv(pos[j.a], rot[j.a], j.rA);
      V<T> pBw = xformv(pos[j.b], rot[j.b], j.rB);
      V<T> C = pAw - pBw;
      if (j.flags & 1u) C = C - j.C0Lin * alpha;
      V<T> F = mulv(j.penaltyLin, C) + j.lambdaLin;
      T js = isA ? T(1) : T(-1);
      V<T> rW = isA ? qrot(rot[j.a], j.rA) : qrot(rot[j.b], j.rB);
      M<T> jAng = skew(-(rW * js)), jAngT = jAng.tr(), K = M<T>::diag(j.penaltyLin), jAngTk = jAngT.mul(K);
      s.ll = s.ll + K;
      s.aa = s.aa + jAngTk.mul(jAng);
      s.cr = s.cr + jAngTk * js;
      V<T> r = rW * js;
      M<T> H = geomBS(0, r) * F.x + geomBS(1, r) * F.y + geomBS(2, r) * F.z;
      V<T> dg;
      for (int c = 0; c < 3; c++) dg[c] = std::sqrt(H.a[0][c] * H.a[0][c] + H.a[1][c] * H.a[1][c] + H.a[2][c] * H.a[2][c]);
      s.aa = s.aa + M<T>::diag(dg);
      s.gl = s.gl + F * js;
      s.ga = s.ga + jAngT.mulv(F);
    }
    if (dotv(j.penaltyAng, j.penaltyAng) > 0) {
      Q<T> qA = j.a == WORLD ? Q<T>{0, 0, 0, 1} : rot[j.a];
      Q<T> qAr = qmul(qA, j.restRel);
      if (j.hinge) {
        V<T> aB = qrot(rot[j.b], j.axis), aA = qrot(qAr, j.axis);
        V<T> C = crossv(aA, aB) * j.torqueArm;
        if (j.flags & 2u) C = C - j.C0Ang * alpha;
        V<T> F = mulv(j.penaltyAng, C) + j.lambdaAng;
        T al = dotv(aA, aB);
        M<T> I3 = M<T>::diag(V<T>(1, 1, 1));
        M<T> Jm = isA ? M<T>::outer(aA, aB) + I3 * (-al) : I3 * al + M<T>::outer(aB, aA) * T(-1);
        Jm = Jm * j.torqueArm;
        M<T> JT = Jm.tr();
        s.aa = s.aa + JT.mul(M<T>::diag(j.penaltyAng)).mul(Jm);
        s.ga = s.ga + JT.mulv(F);
        if (j.motorTorque > 0 || j.limLo < j.limHi) {
          T tw = twistOf(qAr, rot[j.b], j.axis);
          T sm = isA ? T(-1) : T(1);
          if (j.armature > 0) {
            T Ca = wrapPi(tw - j.predTwist);
            T ka = j.armature / std::max(dt * dt, T(1e-12));
            s.aa = s.aa + M<T>::outer(aB, aB) * ka;
            s.ga = s.ga + aB * (ka * Ca * sm);
          }
          if (j.limLo < j.limHi) {
            T over = std::max(tw - j.limHi, T(0)) + std::min(tw - j.limLo, T(0));
            if (over != 0) {
              T kL = 4.0e4;
              T FL = clampT(kL * over, T(-3000), T(3000));
              s.aa = s.aa + M<T>::outer(aB, aB) * kL;
              s.ga = s.ga + aB * (FL * sm);
            }
          }
          if (j.motorTorque > 0) {
            uint32_t mode = j.flags & (3u << 4);
            if (mode == (2u << 4) || mode == (3u << 4)) {
              s.ga = s.ga + aB * (-j.effort * sm);
            } else {
              T Cm = wrapPi(tw - j.motorTarget);
              Q<T> qA0 = j.a == WORLD ? Q<T>{0, 0, 0, 1} : initAng[j.a];
              T tw0 = twistOf(qmul(qA0, j.restRel), initAng[j.b], j.axis);
              T dT = wrapPi(tw - tw0);
              T cD = j.kd / std::max(dt, T(1e-6));
              T raw = j.kp * Cm + cD * dT;
              T eff = clampT(raw, -j.motorTorque, j.motorTorque);
              if (std::fabs(raw) < j.motorTorque) s.aa = s.aa + M<T>::outer(aB, aB) * (j.kp + cD);
              s.ga = s.ga + aB * (eff * sm);
            }
          }
        }
      } else {
        V<T> C = qsub(qAr, rot[j.b]) * j.torqueArm;
        T sg = (isA ? T(1) : T(-1)) * j.torqueArm;
        if (j.flags & 2u) C = C - j.C0Ang * alpha;
        V<T> F = mulv(j.penaltyAng, C) + j.lambdaAng;
        s.aa = s.aa + M<T>::diag(j.penaltyAng * (sg * sg));
        s.ga = s.ga + F * sg;
      }
    }
  }

  void stampSpring(const Spring<T>& sp, uint32_t self, Sys& s) const {
    V<T> pAw = xformv(pos[sp.a], rot[sp.a], sp.rA), pBw = xformv(pos[sp.b], rot[sp.b], sp.rB);
    V<T> d = pAw - pBw;
    T L = lenv(d);
    if (L <= T(1e-6)) return;
    V<T> n = d / L;
    T k = sp.stiffness, f;
    if (sp.hard) {
      T C = L - sp.rest - sp.C0 * (T)set.alpha;
      k = sp.penalty;
      f = k * C + sp.lambda;
      if (f <= 0) return;
    } else {
      f = k * (L - sp.rest);
    }
    bool isA = self == sp.a;
    V<T> rW = isA ? qrot(rot[sp.a], sp.rA) : qrot(rot[sp.b], sp.rB);
    V<T> jl = isA ? n : -n;
    V<T> ja = isA ? crossv(rW, n) : -crossv(rW, n);
    s.ll = s.ll + M<T>::outer(jl, jl) * k;
    s.aa = s.aa + M<T>::outer(ja, ja) * k;
    s.cr = s.cr + M<T>::outer(ja, jl) * k;
    s.gl = s.gl + jl * f;
    s.ga = s.ga + ja * f;
    T ft = std::max(f, T(0)) / L;
    if (ft > 0) s.ll = s.ll + M<T>::diag(V<T>(ft, ft, ft)) + M<T>::outer(n, n) * (-ft);
  }


  // -------------------------------------------------------------------------
  // cloth contacts: rigid feature vs triangle (25_cloth.metal:928-1142), the
  // Planar-DAT pair stream (27_planar_dat.metal:35-229) and OGC emission
  // (27_planar_dat.metal:344-510, 25_cloth.metal:273-347).  The pair set is
  // computed on exact cells (the kernels' hash buckets only add candidates
  // that the exact cell-range / ownership tests reject), each producer's
  // partners ascending.
  // -------------------------------------------------------------------------
  static V<T> closestTri(const V<T>& p, const V<T>& a, const V<T>& b, const V<T>& c, V<T>& bary) {
    V<T> ab = b - a, ac = c - a, ap = p - a;
    T d1 = dotv(ab, ap), d2 = dotv(ac, ap);
    if (d1 <= 0 && d2 <= 0) { bary = V<T>(1, 0, 0); return a; }
    V<T> bpv = p - b;
    T d3 = dotv(ab, bpv), d4 = dotv(ac, bpv);
    if (d3 >= 0 && d4 <= d3) { bary = V<T>(0, 1, 0); return b; }
    T vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      T v = d1 / std::max(d1 - d3, T(1e-12));
      bary = V<T>(1 - v, v, 0);
      return a + ab * v;
    }
    V<T> cp = p - c;
    T d5 = dotv(ab, cp), d6 = dotv(ac, cp);
    if (d6 >= 0 && d5 <= d6) { bary = V<T>(0, 0, 1); return c; }
    T vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
      T w = d2 / std::max(d2 - d6, T(1e-12));
      bary = V<T>(1 - w, 0, w);
      return a + ac * w;
    }
    T va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
      T w = (d4 - d3) / std::max((d4 - d3) + (d5 - d6), T(1e-12));
      bary = V<T>(0, 1 - w, w);
      return b + (c - b) * w;
    }
    T den = va + vb + vc;
    if (std::fabs(den) < T(1e-20)) { bary = V<T>(1, 0, 0); return a; }
    T v = vb / den, w = vc / den;
    bary = V<T>(1 - v - w, v, w);
    return a + ab * v + ac * w;
  }
  static void segSeg(const V<T>& p0, const V<T>& p1, const V<T>& q0, const V<T>& q1, T& s_, T& t_) {
    V<T> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    T a = dotv(d1, d1), e = dotv(d2, d2), f = dotv(d2, r);
    s_ = 0; t_ = 0;
    if (a <= T(1e-12) && e <= T(1e-12)) return;
    if (a <= T(1e-12)) { t_ = clampT(f / e, T(0), T(1)); return; }
    T c = dotv(d1, r);
    if (e <= T(1e-12)) { s_ = clampT(-c / a, T(0), T(1)); return; }
    T b = dotv(d1, d2), den = a * e - b * b;
    if (std::fabs(den) > T(1e-12)) s_ = clampT((b * f - c * e) / den, T(0), T(1));
    t_ = (b * s_ + f) / e;
    if (t_ < 0) { t_ = 0; s_ = clampT(-c / a, T(0), T(1)); }
    else if (t_ > 1) { t_ = 1; s_ = clampT((b - c) / a, T(0), T(1)); }
  }
  static void orthoBasis(const V<T>& n, V<T>& t1, V<T>& t2) {
    // 00_common.metal:224-228
    t1 = std::fabs(n.x) > std::fabs(n.z) ? V<T>(-n.y, n.x, 0) : V<T>(0, -n.z, n.y);
    t1 = normv(t1);
    t2 = crossv(n, t1);
  }
  bool nbrHas(uint32_t v, uint32_t x) const {
    auto b = nbrList.begin() + nbrStart[v], e = b + nbrCount[v];
    return std::binary_search(b, e, x);
  }
  struct Cell { int x, y, z; bool operator<(const Cell& o) const { return std::tie(x, y, z) < std::tie(o.x, o.y, o.z); } };
  Cell elemCell(const V<T>& p) const {
    T cs = (T)elemCellSize;
    return Cell{(int)std::floor(p.x / cs), (int)std::floor(p.y / cs), (int)std::floor(p.z / cs)};
  }
  static Cell clampCell(Cell raw, Cell c0) {
    return Cell{std::min(raw.x, c0.x + 3), std::min(raw.y, c0.y + 3), std::min(raw.z, c0.z + 3)};
  }
  static bool finite3(const V<T>& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
  void elemBounds(uint32_t g, V<T>& lo, V<T>& hi) const {
    T pad = (T)datQueryRadius;
    auto vmin = [](const V<T>& a, const V<T>& b) { return V<T>(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); };
    auto vmax = [](const V<T>& a, const V<T>& b) { return V<T>(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); };
    if (g < nTris) {
      const auto& t = tris[g];
      lo = vmin(vmin(pos[t[0]], pos[t[1]]), pos[t[2]]) - V<T>(pad, pad, pad);
      hi = vmax(vmax(pos[t[0]], pos[t[1]]), pos[t[2]]) + V<T>(pad, pad, pad);
    } else {
      const auto& e = edges[g - nTris];
      lo = vmin(pos[e.first], pos[e.second]) - V<T>(pad, pad, pad);
      hi = vmax(pos[e.first], pos[e.second]) + V<T>(pad, pad, pad);
    }
  }
  int softFind(uint32_t keyA, uint32_t keyB) const {
    auto it = softMap.find({keyA, keyB});
    return it == softMap.end() ? -1 : it->second;
  }
  Soft makeSoft(const uint32_t ids[4], const T w[4], const V<T>& n, T friction, const V<T>& anchor, T C0n, T restT,
                T minMass, uint32_t kind, bool rigidA, bool roundA, bool sideNeg, uint32_t keyA, uint32_t keyB) const {
    Soft r;
    V<T> lam, pen;
    int prev = softFind(keyA, keyB);
    if (prev >= 0) {
      const Soft& ps = prevSoft[prev];
      lam = ps.lambda * ((T)set.alpha * (T)set.gamma);
      pen = ps.penalty;
      V<T> t1o, t2o, t1n, t2n;
      orthoBasis(ps.n, t1o, t2o);
      orthoBasis(n, t1n, t2n);
      V<T> lt = t1o * lam.y + t2o * lam.z;
      lam.y = dotv(t1n, lt);
      lam.z = dotv(t2n, lt);
    }
    T dt = set.dt;
    T floorK = std::max(T(1), minMass / (dt * dt));
    for (int k = 0; k < 3; k++) pen[k] = clampT(pen[k] * (T)set.gamma, floorK, T(1e10));
    for (int k = 0; k < 4; k++) { r.ids[k] = ids[k]; r.w[k] = w[k]; }
    r.n = n;
    r.friction = friction;
    r.anchor = rigidA ? anchor : V<T>(restT, 0, 0);
    r.flags = (rigidA ? 1u : 0u) | (sideNeg ? 2u : 0u) | (kind << 2) | (roundA ? 32u : 0u);
    r.C0 = C0n;
    r.lamCap = std::min((T)set.lambda_max, std::max(T(10), T(1e5) * minMass));
    r.lambda = lam;
    r.penalty = pen;
    r.keyA = keyA;
    r.keyB = keyB;
    return r;
  }
  T prevSide(uint32_t keyA, uint32_t keyB, T proposed) const {
    int prev = softFind(keyA, keyB);
    if (prev < 0) return proposed;
    return (prevSoft[prev].flags & 2u) ? T(-1) : T(1);
  }

  void rtEmit() {
    const T big = std::numeric_limits<T>::max();
    T margin = (T)set.collision_margin;
    for (uint32_t gid = 0; gid < nTris; gid++) {
      const auto& tid = tris[gid];
      V<T> a = pos[tid[0]], b = pos[tid[1]], c = pos[tid[2]];
      V<T> m = (a + b + c) / T(3);
      T rt_ = std::max(std::fabs(radiusSigned[tid[0]]), std::max(std::fabs(radiusSigned[tid[1]]), std::fabs(radiusSigned[tid[2]])));
      T rT = std::max(lenv(m - a), std::max(lenv(m - b), lenv(m - c))) + rt_;
      V<T> triN = crossv(b - a, c - a);
      T tl = lenv(triN);
      if (tl < T(1e-12)) continue;
      triN = triN / tl;
      V<T> velT = (vel[tid[0]] + vel[tid[1]] + vel[tid[2]]) / T(3);
      uint32_t emitted = 0;
      auto test = [&](uint32_t col) {
        uint32_t o = cOwner[col];
        V<T> bp = pos[o];
        Q<T> bq = rot[o];
        V<T> cen = bp + qrot(bq, cLocal[col]);
        Q<T> cr = qmul(bq, cLocalRot[col]);
        uint32_t kind = cFlags[col] & 0xFu;
        bool nearO = lenv(cen - m) <= rT + std::fabs(cRad[col]) + margin;
        if (nearO && kind == 0) {
          V<T> lm = qrot(qconj(cr), m - cen);
          V<T> dq2(std::max(std::fabs(lm.x) - cSize[col].x * T(0.5), T(0)),
                   std::max(std::fabs(lm.y) - cSize[col].y * T(0.5), T(0)),
                   std::max(std::fabs(lm.z) - cSize[col].z * T(0.5), T(0)));
          nearO = lenv(dq2) <= rT + margin + T(0.05);
        }
        if (!nearO) return;
        for (int k = 0; k < 3; k++)
          if (excluded(std::min(o, tid[k]), std::max(o, tid[k]))) return;
        V<T> fw[4];
        T fr[4];
        uint32_t fid[4];
        int nf = 0;
        if (kind == 1) { fw[0] = cen; fr[0] = cSize[col].x * T(0.5); fid[0] = 8; nf = 1; }
        else if (kind == 3) {
          // capsule: spheres at its two ends and middle (25_cloth.metal:957-970)
          T half = cSize[col].x * T(0.5), rc = cSize[col].y;
          V<T> ax = qrot(cr, V<T>(0, 0, 1));
          for (int i = -1; i <= 1; i++) {
            V<T> w = cen + ax * (half * T(i));
            if (lenv(w - m) > rT + rc + T(4) * margin) continue;
            fw[nf] = w; fr[nf] = rc; fid[nf] = (uint32_t)(10 + i); nf++;
          }
        } else if (kind == 0) {
          V<T> h3 = cSize[col] * T(0.5);
          for (uint32_t i = 0; i < 8; i++) {
            V<T> loc((i & 1) ? h3.x : -h3.x, (i & 2) ? h3.y : -h3.y, (i & 4) ? h3.z : -h3.z);
            V<T> w = xformv(cen, cr, loc);
            if (lenv(w - m) > rT + T(4) * margin) continue;
            if (nf < 4) { fw[nf] = w; fr[nf] = 0; fid[nf] = i; nf++; }
          }
        }
        for (int fi = 0; fi < nf && emitted < 4; fi++) {
          V<T> bary;
          V<T> q = closestTri(fw[fi], a, b, c, bary);
          V<T> dv = fw[fi] - q;
          T dist = lenv(dv), hSum = rt_ + fr[fi];
          T inflate = std::min(lenv(vel[o] - velT) * (T)set.dt, T(0.5) * rT);
          if (dist - hSum > (T)elemMargin + inflate) continue;
          T side = dotv(dv, triN) >= 0 ? T(1) : T(-1);
          uint32_t keyA = (2u << 28) | col, keyB = gid * 16u + fid[fi];
          T proposed = dotv(bp - q, triN) >= 0 ? T(1) : T(-1);
          bool interior = bary.x > T(0.02) && bary.y > T(0.02) && bary.z > T(0.02);
          T sMem = interior ? prevSide(keyA, keyB, proposed) : proposed;
          V<T> n;
          T g;
          if (dist > T(1e-7) && side == sMem) { n = dv / dist; g = dist; }
          else { n = triN * sMem; g = -dist; }
          T fricT = (friction[tid[0]] + friction[tid[1]] + friction[tid[2]]) / T(3);
          T fric = std::sqrt(std::max(friction[o] * fricT, T(0)));
          T minMass = mass[o] > 0 ? mass[o] : big;
          for (int k = 0; k < 3; k++) if (mass[tid[k]] > 0) minMass = std::min(minMass, mass[tid[k]]);
          if (minMass == big) continue;
          bool roundA = kind != 0;
          V<T> anchor = roundA ? fw[fi] - bp : qrot(qconj(bq), fw[fi] - bp);
          uint32_t ids[4] = {o, tid[0], tid[1], tid[2]};
          T w[4] = {T(1), -bary.x, -bary.y, -bary.z};
          Soft r = makeSoft(ids, w, n, fric, anchor, g - hSum + (T)elemMargin, hSum - (T)elemMargin, minMass, 2u, true,
                            roundA, sMem < 0, keyA, keyB);
          softCandidates++;
          if (soft.size() < maxSoft) soft.push_back(r);
          emitted++;
        }
      };
      if (nHashedRigid > 0 && nHashedRigid <= 8) {
        for (uint32_t ri = 0; ri < nHashedRigid && emitted < 4; ri++) test(hashedRigidList[ri]);
      } else if (nHashedRigid > 0) {
        int R = (int)std::ceil((rT + T(0.5) * (T)cellSize) / (T)cellSize);
        R = std::max(1, std::min(4, R));
        int cx, cy, cz;
        cellOf(m, cx, cy, cz);
        for (int dz = -R; dz <= R && emitted < 4; dz++)
          for (int dy = -R; dy <= R && emitted < 4; dy++)
            for (int dx = -R; dx <= R && emitted < 4; dx++) {
              uint32_t h = cellHash32(cx + dx, cy + dy, cz + dz, 0u, hashSize);
              for (uint32_t k = cellStart[h]; k < cellStart[h] + cellCount[h] && emitted < 4; k++) {
                uint32_t col = cellBodies[k];
                if (cRad[col] < 0) continue;
                test(col);
              }
            }
      }
      for (uint32_t gI = 0; gI < globals.size() && emitted < 4; gI++) {
        uint32_t col = globals[gI];
        if (cRad[col] < 0) continue;
        test(col);
      }
    }
    softRT = softCandidates;
  }

  void datBuild() {
    datPairs.clear();
    datTotal = datVT = 0;
    std::map<Cell, std::vector<uint32_t>> grid;
    for (uint32_t g = 0; g < nTris + (uint32_t)edges.size(); g++) {
      V<T> lo, hi;
      elemBounds(g, lo, hi);
      if (!finite3(lo) || !finite3(hi)) { datNonfinite++; continue; }
      Cell c0 = elemCell(lo), raw = elemCell(hi);
      if (raw.x - c0.x > 3 || raw.y - c0.y > 3 || raw.z - c0.z > 3) datOverflow++;
      Cell c1 = clampCell(raw, c0);
      for (int z = c0.z; z <= c1.z; z++)
        for (int y = c0.y; y <= c1.y; y++)
          for (int x = c0.x; x <= c1.x; x++) grid[Cell{x, y, z}].push_back(g);
    }
    T q2 = (T)datQueryRadius * (T)datQueryRadius;
    auto emit = [&](uint32_t kind, uint32_t owner, std::vector<uint32_t>& others) {
      std::sort(others.begin(), others.end());
      for (uint32_t o : others) {
        datTotal++;
        if (datPairs.size() < maxDatPairs) datPairs.push_back({kind, owner, o});
      }
    };
    for (uint32_t v : particles) {
      V<T> p = pos[v];
      if (!finite3(p)) { datNonfinite++; continue; }
      Cell cell = elemCell(p);
      std::vector<uint32_t> acc;
      auto it = grid.find(cell);
      if (it != grid.end()) {
        for (uint32_t e : it->second) {
          if (e >= nTris) continue;
          const auto& t = tris[e];
          if (t[0] == v || t[1] == v || t[2] == v) continue;
          uint32_t g = clothGroup[v];
          bool same = g != 0 && clothGroup[t[0]] == g && clothGroup[t[1]] == g && clothGroup[t[2]] == g;
          if (same && !selfColl[v]) continue;
          if (same && (nbrHas(v, t[0]) || nbrHas(v, t[1]) || nbrHas(v, t[2]))) continue;
          V<T> bary;
          V<T> q = closestTri(p, pos[t[0]], pos[t[1]], pos[t[2]], bary);
          V<T> dv = p - q;
          if (dotv(dv, dv) > q2) continue;
          acc.push_back(e);
        }
      }
      emit(0u, v, acc);
    }
    datVT = datTotal;
    for (uint32_t gid = 0; gid < edges.size(); gid++) {
      auto a = edges[gid];
      V<T> a0 = pos[a.first], a1 = pos[a.second];
      if (!finite3(a0) || !finite3(a1)) { datNonfinite++; continue; }
      auto vmin = V<T>(std::min(a0.x, a1.x), std::min(a0.y, a1.y), std::min(a0.z, a1.z));
      auto vmax = V<T>(std::max(a0.x, a1.x), std::max(a0.y, a1.y), std::max(a0.z, a1.z));
      Cell ac0 = elemCell(vmin), raw = elemCell(vmax);
      if (raw.x - ac0.x > 3 || raw.y - ac0.y > 3 || raw.z - ac0.z > 3) datOverflow++;
      Cell ac1 = clampCell(raw, ac0);
      std::vector<uint32_t> acc;
      for (int z = ac0.z; z <= ac1.z; z++)
        for (int y = ac0.y; y <= ac1.y; y++)
          for (int x = ac0.x; x <= ac1.x; x++) {
            auto it = grid.find(Cell{x, y, z});
            if (it == grid.end()) continue;
            for (uint32_t e : it->second) {
              if (e < nTris) continue;
              uint32_t other = e - nTris;
              if (other <= gid) continue;
              auto b = edges[other];
              if (b.first == a.first || b.first == a.second || b.second == a.first || b.second == a.second) continue;
              uint32_t g = clothGroup[a.first];
              bool same = g != 0 && clothGroup[a.second] == g && clothGroup[b.first] == g && clothGroup[b.second] == g;
              if (same && !selfColl[a.first]) continue;
              if (same && (nbrHas(a.first, b.first) || nbrHas(a.first, b.second) || nbrHas(a.second, b.first) ||
                           nbrHas(a.second, b.second)))
                continue;
              V<T> blo, bhi;
              elemBounds(nTris + other, blo, bhi);
              Cell bc0 = elemCell(blo), bc1 = clampCell(elemCell(bhi), bc0);
              Cell o0{std::max(ac0.x, bc0.x), std::max(ac0.y, bc0.y), std::max(ac0.z, bc0.z)};
              Cell o1{std::min(ac1.x, bc1.x), std::min(ac1.y, bc1.y), std::min(ac1.z, bc1.z)};
              if (o0.x > o1.x || o0.y > o1.y || o0.z > o1.z) continue;
              if (x != o0.x || y != o0.y || z != o0.z) continue;
              T s_, t_;
              V<T> b0 = pos[b.first], b1 = pos[b.second];
              segSeg(a0, a1, b0, b1, s_, t_);
              V<T> dv = (a0 + (a1 - a0) * s_) - (b0 + (b1 - b0) * t_);
              if (dotv(dv, dv) > q2) continue;
              acc.push_back(other);
            }
          }
      emit(1u, gid, acc);
    }
    // the frame reports the larger demand of its two queries (dat_finalize_pairs,
    // 27_planar_dat.metal:208-229, atomic max)
    datPeak = std::max(datPeak, datTotal);
    datPeakVT = std::max(datPeakVT, datVT);
  }

  bool datContact(const std::array<uint32_t, 3>& pr, Soft& out) const {
    const T big = std::numeric_limits<T>::max();
    const std::vector<V<T>>& ref = initLin;
    T rq = (T)datQueryRadius, em = (T)elemMargin;
    if (pr[0] == 0) {
      uint32_t v = pr[1], ti = pr[2];
      const auto& t = tris[ti];
      V<T> p = pos[v], a = pos[t[0]], b = pos[t[1]], c = pos[t[2]];
      V<T> bary;
      V<T> q = closestTri(p, a, b, c, bary);
      V<T> dv = p - q;
      T dist2 = dotv(dv, dv);
      T rv = std::fabs(radiusSigned[v]);
      T rt = std::max(std::fabs(radiusSigned[t[0]]), std::max(std::fabs(radiusSigned[t[1]]), std::fabs(radiusSigned[t[2]])));
      T hSum = rv + rt;
      V<T> velT = (vel[t[0]] + vel[t[1]] + vel[t[2]]) / T(3);
      T inflate = std::min(lenv(vel[v] - velT) * (T)set.dt, T(0.3) * (T)surfaceContactCellSize);
      T detect = std::min(hSum + em + inflate, rq);
      if (dist2 > detect * detect) return false;
      T dist = std::sqrt(dist2);
      V<T> triN = crossv(b - a, c - a);
      T nl = lenv(triN);
      if (nl < T(1e-8)) return false;
      triN = triN / nl;
      T side = dotv(dv, triN) >= 0 ? T(1) : T(-1);
      uint32_t keyA = (1u << 28) | v, keyB = ti;
      bool interior = bary.x > T(0.02) && bary.y > T(0.02) && bary.z > T(0.02);
      T rem = interior ? prevSide(keyA, keyB, side) : side;
      V<T> n;
      T gap;
      if (dist > T(1e-7) && side == rem) { n = dv / dist; gap = dist; }
      else { n = triN * rem; gap = -dist; }
      T fricT = friction[t[0]] * bary.x + friction[t[1]] * bary.y + friction[t[2]] * bary.z;
      T fric = std::sqrt(std::max(friction[v] * fricT, T(0)));
      T minMass = mass[v] > 0 ? mass[v] : big;
      for (int j = 0; j < 3; j++) if (mass[t[j]] > 0) minMass = std::min(minMass, mass[t[j]]);
      if (minMass == big) return false;
      V<T> dq = (p - ref[v]) - (a - ref[t[0]]) * bary.x - (b - ref[t[1]]) * bary.y - (c - ref[t[2]]) * bary.z;
      T C0n = (gap - hSum + em) - dotv(n, dq);
      uint32_t ids[4] = {v, t[0], t[1], t[2]};
      T w[4] = {T(1), -bary.x, -bary.y, -bary.z};
      out = makeSoft(ids, w, n, fric, V<T>(), C0n, hSum - em, minMass, 1u, false, false, rem < 0, keyA, keyB);
      return true;
    }
    uint32_t ea = pr[1], eb = pr[2];
    auto ai = edges[ea], bi = edges[eb];
    // contacts between authored edges or within a self-colliding surface;
    // tet-boundary edge pairs only bound the motion (27_planar_dat.metal:196-201)
    if (ea >= numContactEdges || eb >= numContactEdges) {
      uint32_t g = clothGroup[ai.first];
      bool same = g != 0 && clothGroup[ai.second] == g && clothGroup[bi.first] == g && clothGroup[bi.second] == g;
      if (!(same && selfColl[ai.first])) return false;
    }
    V<T> a0 = pos[ai.first], a1 = pos[ai.second], b0 = pos[bi.first], b1 = pos[bi.second];
    T s_, t_;
    segSeg(a0, a1, b0, b1, s_, t_);
    const T eps = T(0.02);
    if (s_ < eps || s_ > T(1) - eps || t_ < eps || t_ > T(1) - eps) return false;
    V<T> dv = (a0 + (a1 - a0) * s_) - (b0 + (b1 - b0) * t_);
    T dist = lenv(dv);
    T rA = std::max(std::fabs(radiusSigned[ai.first]), std::fabs(radiusSigned[ai.second]));
    T rB = std::max(std::fabs(radiusSigned[bi.first]), std::fabs(radiusSigned[bi.second]));
    T hSum = rA + rB;
    V<T> va = (vel[ai.first] + vel[ai.second]) * T(0.5), vb = (vel[bi.first] + vel[bi.second]) * T(0.5);
    T inflate = std::min(lenv(va - vb) * (T)set.dt, T(0.3) * (T)surfaceContactCellSize);
    T detect = std::min(hSum + em + inflate, rq);
    if (dist > detect) return false;
    uint32_t keyA = (3u << 28) | ea, keyB = eb;
    int prev = softFind(keyA, keyB);
    V<T> n;
    T gap = dist;
    if (std::isfinite(dist) && dist >= T(1e-7)) {
      V<T> nGeo = dv / dist;
      n = nGeo;
      if (prev >= 0 && dotv(nGeo, prevSoft[prev].n) < 0) { n = -nGeo; gap = -dist; }
    } else if (prev >= 0 && finite3(prevSoft[prev].n)) {
      n = prevSoft[prev].n;
      gap = 0;
    } else {
      V<T> ua = a1 - a0, ub = b1 - b0, rel = va - vb;
      V<T> raw = crossv(ua, ub);
      T mag = lenv(raw);
      if (std::isfinite(mag) && mag >= T(1e-8)) {
        n = raw / mag;
      } else {
        T la = lenv(ua), lb = lenv(ub);
        V<T> u = la >= lb ? ua : ub;
        T ul = std::max(la, lb);
        if (!std::isfinite(ul) || ul < T(1e-8)) return false;
        u = u / ul;
        V<T> perp = rel - u * dotv(u, rel);
        T pl = lenv(perp);
        if (!std::isfinite(pl) || pl < T(1e-8)) {
          V<T> axis = std::fabs(u.x) < T(0.7) ? V<T>(1, 0, 0) : V<T>(0, 1, 0);
          perp = crossv(u, axis);
          pl = lenv(perp);
        }
        if (!std::isfinite(pl) || pl < T(1e-8)) return false;
        n = perp / pl;
      }
      if (dotv(n, rel) > 0) n = -n;
      if (!finite3(n)) return false;
      gap = 0;
    }
    T fric = std::sqrt(std::max((friction[ai.first] + friction[ai.second]) * T(0.5) *
                                (friction[bi.first] + friction[bi.second]) * T(0.5), T(0)));
    T minMass = big;
    uint32_t ids[4] = {ai.first, ai.second, bi.first, bi.second};
    for (int j = 0; j < 4; j++) if (mass[ids[j]] > 0) minMass = std::min(minMass, mass[ids[j]]);
    if (minMass == big) return false;
    V<T> dq = (a0 - ref[ai.first]) * (T(1) - s_) + (a1 - ref[ai.second]) * s_ - (b0 - ref[bi.first]) * (T(1) - t_) -
              (b1 - ref[bi.second]) * t_;
    T C0n = (gap - hSum + em) - dotv(n, dq);
    T w[4] = {T(1) - s_, s_, -(T(1) - t_), -t_};
    out = makeSoft(ids, w, n, fric, V<T>(), C0n, hSum - em, minMass, 3u, false, false, false, keyA, keyB);
    return true;
  }

  // -------------------------------------------------------------------------
  // isotropic OGC mode (25_cloth.metal:350-940, 40_solver.metal:697-717,
  // 1361-1382, 1676-1681; order GPUSolver.swift:3686-3756, 4243-4275).
  // Each surface particle keeps its four closest collision triangles across
  // frames and each authored edge its four closest edges; per frame a small
  // candidate pool is re-evaluated -- the tracked set, the triangles adjacent
  // to the best, the mesh neighbours' best -- and the hashed element grid is
  // scanned on staggered reseeds.  Every observed vertex-facet / edge-edge
  // distance outside the 2-ring lowers the squared clearance bound of the
  // vertices involved; the predictor and every primal step keep a particle
  // within max(0.45 sqrt(bound), 0.2 r) of the pose the bounds hold around,
  // and a refresh re-anchors that pose once more than 1 % of the particles
  // were held back.  Contacts are emitted from the tracked sets.
  // Deterministic where the reference is not: every tracker reads the
  // previous frame's tracked sets (the reference reads its neighbours' as
  // they are being rewritten) and ties in distance keep the lower element id.
  // -------------------------------------------------------------------------
  static constexpr double kOgcFar = 3.3961514e38;  // the bounds' clear value (f32 bits 0x7F7F7F7F)
  struct Best4 {
    uint32_t id[4];
    T d2[4];
  };
  static void b4Init(Best4& b) {
    for (int i = 0; i < 4; i++) { b.id[i] = 0xFFFFFFFFu; b.d2[i] = (T)std::numeric_limits<float>::max(); }
  }
  static bool b4Has(const Best4& b, uint32_t t) { return b.id[0] == t || b.id[1] == t || b.id[2] == t || b.id[3] == t; }
  // the kept set is the four smallest (d2, id); a NaN distance never enters
  static void b4Insert(Best4& b, uint32_t t, T d2) {
    if (!(d2 < b.d2[3] || (d2 == b.d2[3] && t < b.id[3]))) return;
    int i = 3;
    for (; i > 0 && (b.d2[i - 1] > d2 || (b.d2[i - 1] == d2 && b.id[i - 1] > t)); i--) {
      b.id[i] = b.id[i - 1];
      b.d2[i] = b.d2[i - 1];
    }
    b.id[i] = t;
    b.d2[i] = d2;
  }
  bool inList(const std::vector<uint32_t>& list, uint32_t s, uint32_t n, uint32_t x) const {
    return std::binary_search(list.begin() + s, list.begin() + s + n, x);
  }
  bool in1(uint32_t v, uint32_t x) const { return inList(nbrList, nbrStart[v], nbrCount[v], x); }
  bool in2(uint32_t v, uint32_t x) const { return inList(nbr2List, nbr2Start[v], nbr2Count[v], x); }
  void ogcLower(uint32_t v, T d2) {
    if (d2 < ogcBound[v]) ogcBound[v] = d2;
  }
  // hashed element grid at the contact reach: triangles, then authored edges
  // (entry | 0x80000000), each binned into at most 3 cells per axis
  std::vector<std::vector<uint32_t>> isoBuckets;
  std::array<int, 3> isoCell(const V<T>& p) const {
    const float cs = surfaceContactCellSize;
    return {(int)std::floor((float)p.x / cs), (int)std::floor((float)p.y / cs), (int)std::floor((float)p.z / cs)};
  }
  void isoGrid() {
    isoBuckets.assign(elemHashSize, {});
    const float em = elemMargin;
    for (uint32_t g = 0; g < nTris + numContactEdges; g++) {
      V<T> lo, hi;
      T pad;
      uint32_t entry;
      if (g < nTris) {
        const auto& t = tris[g];
        pad = (T)(std::max(std::fabs((float)radiusSigned[t[0]]),
                           std::max(std::fabs((float)radiusSigned[t[1]]), std::fabs((float)radiusSigned[t[2]]))) + em);
        lo = V<T>(std::min(std::min(pos[t[0]].x, pos[t[1]].x), pos[t[2]].x),
                  std::min(std::min(pos[t[0]].y, pos[t[1]].y), pos[t[2]].y),
                  std::min(std::min(pos[t[0]].z, pos[t[1]].z), pos[t[2]].z));
        hi = V<T>(std::max(std::max(pos[t[0]].x, pos[t[1]].x), pos[t[2]].x),
                  std::max(std::max(pos[t[0]].y, pos[t[1]].y), pos[t[2]].y),
                  std::max(std::max(pos[t[0]].z, pos[t[1]].z), pos[t[2]].z));
        entry = g;
      } else {
        const auto& e = edges[g - nTris];
        pad = (T)(std::max(std::fabs((float)radiusSigned[e.first]), std::fabs((float)radiusSigned[e.second])) + em);
        lo = V<T>(std::min(pos[e.first].x, pos[e.second].x), std::min(pos[e.first].y, pos[e.second].y),
                  std::min(pos[e.first].z, pos[e.second].z));
        hi = V<T>(std::max(pos[e.first].x, pos[e.second].x), std::max(pos[e.first].y, pos[e.second].y),
                  std::max(pos[e.first].z, pos[e.second].z));
        entry = 0x80000000u | (g - nTris);
      }
      lo = lo - V<T>(pad, pad, pad);
      hi = hi + V<T>(pad, pad, pad);
      if (!finite3(lo) || !finite3(hi)) continue;  // a non-finite element registers nowhere
      auto c0 = isoCell(lo), c1 = isoCell(hi);
      for (int k = 0; k < 3; k++) c1[k] = std::min(c1[k], c0[k] + 2);
      std::vector<uint32_t> seen;
      for (int z = c0[2]; z <= c1[2]; z++)
        for (int y = c0[1]; y <= c1[1]; y++)
          for (int x = c0[0]; x <= c1[0]; x++) {
            uint32_t h = cellHash32(x, y, z, 0u, elemHashSize);
            if (std::find(seen.begin(), seen.end(), h) != seen.end()) continue;
            seen.push_back(h);
            isoBuckets[h].push_back(entry);
          }
    }
  }
  bool isoReseed(uint32_t id, const V<T>& relV, bool missing, bool solid) const {
    T vth = T(0.25f * surfaceContactCellSize / std::max((float)set.dt, 1e-6f));
    uint32_t period = dotv(relV, relV) > vth * vth ? 1u : 7u;
    return ((id + isoFrame) & period) == 0u || (missing && !solid) || (missing && solid && ((id + isoFrame) & 3u) == 0u);
  }
  uint32_t isoFrame = 0;

  void isoVertex(uint32_t v) {
    const V<T> pv = pos[v];
    const uint32_t gv = clothGroup[v];
    const bool solidV = gv != 0 && clothVert[v] == 0;
    const T cap2 = (T)surfaceContactCellSize * (T)surfaceContactCellSize;
    Best4 best;
    b4Init(best);
    T own = (T)kOgcFar;
    auto consider = [&](uint32_t t) {
      if (t >= nTris || b4Has(best, t)) return;
      const auto& tv = tris[t];
      if (tv[0] == v || tv[1] == v || tv[2] == v) return;
      if (solidV && clothGroup[tv[0]] == gv && clothGroup[tv[1]] == gv && clothGroup[tv[2]] == gv &&
          clothVert[tv[0]] == 0 && clothVert[tv[1]] == 0 && clothVert[tv[2]] == 0)
        return;
      V<T> bary;
      V<T> q = closestTri(pv, pos[tv[0]], pos[tv[1]], pos[tv[2]], bary);
      V<T> dd = pv - q;
      T d2 = dotv(dd, dd);
      if (!in1(v, tv[0]) && !in1(v, tv[1]) && !in1(v, tv[2])) b4Insert(best, t, d2);
      if (d2 < cap2 && !in2(v, tv[0]) && !in2(v, tv[1]) && !in2(v, tv[2])) {
        own = std::min(own, d2);
        ogcLower(tv[0], d2);
        ogcLower(tv[1], d2);
        ogcLower(tv[2], d2);
      }
    };
    const auto& tracked = vtTrackPrev[v];
    for (int i = 0; i < 4; i++) consider(tracked[i]);
    if (tracked[0] < nTris)
      for (int i = 0; i < 3; i++) consider(triAdj[tracked[0]][i]);
    for (uint32_t k = 0; k < nbrCount[v] && k < 8; k++) consider(vtTrackPrev[nbrList[nbrStart[v] + k]][0]);
    V<T> relV = vel[v];
    if (best.id[0] < nTris) {
      const auto& bt = tris[best.id[0]];
      relV = relV - (vel[bt[0]] + vel[bt[1]] + vel[bt[2]]) / T(3);
    }
    if (isoReseed(v, relV, best.id[0] == 0xFFFFFFFFu, solidV)) {
      auto cc = isoCell(pv);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++)
            for (uint32_t entry : isoBuckets[cellHash32(cc[0] + dx, cc[1] + dy, cc[2] + dz, 0u, elemHashSize)])
              if (!(entry & 0x80000000u)) consider(entry);
    }
    vtTrack[v] = {best.id[0], best.id[1], best.id[2], best.id[3]};
    if (own < (T)kOgcFar) ogcLower(v, own);
    // emission from the tracked set (25_cloth.metal:532-629)
    const T big = std::numeric_limits<T>::max(), em = (T)elemMargin;
    const T rv = std::fabs(radiusSigned[v]);
    for (int bi = 0; bi < 4; bi++) {
      uint32_t t = best.id[bi];
      if (t >= nTris) break;
      const auto& tv = tris[t];
      V<T> a = pos[tv[0]], b = pos[tv[1]], c = pos[tv[2]];
      V<T> bary;
      V<T> q = closestTri(pv, a, b, c, bary);
      V<T> dd = pv - q;
      T dist2 = dotv(dd, dd);
      T rt = std::max(std::fabs(radiusSigned[tv[0]]), std::max(std::fabs(radiusSigned[tv[1]]), std::fabs(radiusSigned[tv[2]])));
      T hSum = rv + rt;
      V<T> velT = (vel[tv[0]] + vel[tv[1]] + vel[tv[2]]) / T(3);
      T inflate = std::min(lenv(vel[v] - velT) * (T)set.dt, T(0.3) * (T)surfaceContactCellSize);
      T detect = hSum + em + inflate;
      if (dist2 > detect * detect) continue;
      T dist = std::sqrt(dist2);
      V<T> triN = crossv(b - a, c - a);
      T tnl = lenv(triN);
      if (tnl < T(1e-12)) continue;
      triN = triN / tnl;
      T side = dotv(dd, triN) >= 0 ? T(1) : T(-1);
      uint32_t keyA = (1u << 28) | v, keyB = t;
      bool deep = bary.x > T(0.02) && bary.y > T(0.02) && bary.z > T(0.02);
      T mem = deep ? prevSide(keyA, keyB, side) : side;
      V<T> n;
      T g;
      if (dist > T(1e-7) && side == mem) { n = dd / dist; g = dist; }
      else { n = triN * mem; g = -dist; }
      T fricT = friction[tv[0]] * bary.x + friction[tv[1]] * bary.y + friction[tv[2]] * bary.z;
      T fric = std::sqrt(std::max(friction[v] * fricT, T(0)));
      T minMass = mass[v] > 0 ? mass[v] : big;
      for (int j = 0; j < 3; j++) if (mass[tv[j]] > 0) minMass = std::min(minMass, mass[tv[j]]);
      if (minMass == big) continue;
      uint32_t ids[4] = {v, tv[0], tv[1], tv[2]};
      T w[4] = {T(1), -bary.x, -bary.y, -bary.z};
      Soft r = makeSoft(ids, w, n, fric, V<T>(), g - hSum + em, hSum - em, minMass, 1u, false, false, mem < 0, keyA,
                        keyB);
      softCandidates++;
      if (soft.size() < maxSoft) soft.push_back(r);
    }
  }

  void isoEdge(uint32_t e) {
    const auto eA = edges[e];
    const V<T> a0 = pos[eA.first], a1 = pos[eA.second];
    const uint32_t gA = clothGroup[eA.first];
    const bool solidA = gA != 0 && clothGroup[eA.second] == gA && clothVert[eA.first] == 0 && clothVert[eA.second] == 0;
    const T cap2 = (T)surfaceContactCellSize * (T)surfaceContactCellSize;
    Best4 best;
    b4Init(best);
    T own = (T)kOgcFar;
    auto consider = [&](uint32_t o) {
      if (o >= numContactEdges || o == e || b4Has(best, o)) return;
      const auto eB = edges[o];
      if (eB.first == eA.first || eB.first == eA.second || eB.second == eA.first || eB.second == eA.second) return;
      if (solidA && clothGroup[eB.first] == gA && clothGroup[eB.second] == gA && clothVert[eB.first] == 0 &&
          clothVert[eB.second] == 0)
        return;
      T s_, t_;
      V<T> b0 = pos[eB.first], b1 = pos[eB.second];
      segSeg(a0, a1, b0, b1, s_, t_);
      V<T> dd = (a0 + (a1 - a0) * s_) - (b0 + (b1 - b0) * t_);
      T d2 = dotv(dd, dd);
      if (!in1(eA.first, eB.first) && !in1(eA.first, eB.second) && !in1(eA.second, eB.first) &&
          !in1(eA.second, eB.second))
        b4Insert(best, o, d2);
      if (d2 < cap2 && !in2(eA.first, eB.first) && !in2(eA.first, eB.second)) {
        own = std::min(own, d2);
        ogcLower(eB.first, d2);
        ogcLower(eB.second, d2);
      }
    };
    const auto& tracked = eeTrackPrev[e];
    for (int i = 0; i < 4; i++) consider(tracked[i]);
    for (int side = 0; side < 2; side++) {
      uint32_t vv = side == 0 ? eA.first : eA.second;
      uint32_t cnt = std::min<uint32_t>(vEdgeCount[vv], 6u);
      for (uint32_t k = 0; k < cnt; k++) {
        uint32_t ie = vEdgeList[vEdgeStart[vv] + k];
        if (ie < numContactEdges) consider(eeTrackPrev[ie][0]);
      }
    }
    V<T> velA = (vel[eA.first] + vel[eA.second]) * T(0.5);
    V<T> relV = velA;
    if (best.id[0] < numContactEdges) {
      const auto bt = edges[best.id[0]];
      relV = relV - (vel[bt.first] + vel[bt.second]) * T(0.5);
    }
    if (isoReseed(e, relV, best.id[0] == 0xFFFFFFFFu, solidA)) {
      auto cc = isoCell((a0 + a1) * T(0.5));
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++)
            for (uint32_t entry : isoBuckets[cellHash32(cc[0] + dx, cc[1] + dy, cc[2] + dz, 0u, elemHashSize)])
              if (entry & 0x80000000u) consider(entry & ~0x80000000u);
    }
    eeTrack[e] = {best.id[0], best.id[1], best.id[2], best.id[3]};
    if (own < (T)kOgcFar) {
      ogcLower(eA.first, own);
      ogcLower(eA.second, own);
    }
    // emission, each pair once by its lower edge (25_cloth.metal:790-853)
    const T big = std::numeric_limits<T>::max(), em = (T)elemMargin;
    const T rA = std::max(std::fabs(radiusSigned[eA.first]), std::fabs(radiusSigned[eA.second]));
    for (int bi = 0; bi < 4; bi++) {
      uint32_t o = best.id[bi];
      if (o >= numContactEdges) break;
      if (o <= e) continue;
      const auto eB = edges[o];
      V<T> b0 = pos[eB.first], b1 = pos[eB.second];
      T s_, t_;
      segSeg(a0, a1, b0, b1, s_, t_);
      if (s_ < T(0.02) || s_ > T(1) - T(0.02) || t_ < T(0.02) || t_ > T(1) - T(0.02)) continue;
      V<T> dd = (a0 + (a1 - a0) * s_) - (b0 + (b1 - b0) * t_);
      T dist = lenv(dd);
      T rB = std::max(std::fabs(radiusSigned[eB.first]), std::fabs(radiusSigned[eB.second]));
      T hSum = rA + rB;
      V<T> velB = (vel[eB.first] + vel[eB.second]) * T(0.5);
      T inflate = std::min(lenv(velA - velB) * (T)set.dt, T(0.3) * (T)surfaceContactCellSize);
      if (dist - hSum > em + inflate) continue;
      if (dist < T(1e-7)) continue;
      V<T> nGeo = dd / dist;
      uint32_t keyA = (3u << 28) | e, keyB = o;
      V<T> n = nGeo;
      T g = dist;
      int prev = softFind(keyA, keyB);
      if (prev >= 0 && dotv(nGeo, prevSoft[prev].n) < 0) { n = -nGeo; g = -dist; }
      T fric = std::sqrt(std::max((friction[eA.first] + friction[eA.second]) * T(0.5) *
                                  (friction[eB.first] + friction[eB.second]) * T(0.5), T(0)));
      T minMass = big;
      uint32_t ids[4] = {eA.first, eA.second, eB.first, eB.second};
      for (int j = 0; j < 4; j++) if (mass[ids[j]] > 0) minMass = std::min(minMass, mass[ids[j]]);
      if (minMass == big) continue;
      T w[4] = {T(1) - s_, s_, -(T(1) - t_), -t_};
      Soft r = makeSoft(ids, w, n, fric, V<T>(), g - hSum + em, hSum - em, minMass, 3u, false, false, false, keyA, keyB);
      softCandidates++;
      if (soft.size() < maxSoft) soft.push_back(r);
    }
  }

  void isoEmit() {
    isoFrame = (uint32_t)(frame - 1);  // the frame index before this step
    ogcBound.assign(nb, (T)kOgcFar);
    ogcExceeded = 0;
    ogcTruncFrame = ogcRefreshFrame = 0;
    std::swap(vtTrack, vtTrackPrev);
    std::swap(eeTrack, eeTrackPrev);
    isoGrid();
    for (uint32_t v : particles) isoVertex(v);
    for (uint32_t e = 0; e < numContactEdges; e++) isoEdge(e);
  }
  // the displacement cap (40_solver.metal:1361-1382): returns the kept dx
  V<T> ogcTruncate(uint32_t body, const V<T>& cur, V<T> dx) {
    T d2 = ogcBound[body];
    if (!(d2 < T(1e37))) return dx;
    T bv = std::max(T(0.45) * std::sqrt(d2), T(0.2) * std::fabs(radiusSigned[body]));
    V<T> tot = (cur + dx) - ogcPrev[body];
    T t2 = dotv(tot, tot);
    if (t2 > bv * bv) {
      tot = tot * (bv / std::sqrt(t2));
      dx = (ogcPrev[body] + tot) - cur;
#pragma omp atomic
      ogcExceeded++;
#pragma omp atomic
      ogcTruncFrame++;
    }
    return dx;
  }
  // Alg 3 refresh (25_cloth.metal:857-939): when more than 1 % of the
  // particles were held back, bounds are recomputed from the tracked sets at
  // the live pose, which becomes the new anchor
  void ogcRefresh() {
    const bool fire = ogcExceeded * 100u > particles.size();
    ogcExceeded = 0;
    if (!fire) return;
    ogcRefreshFrame++;
    for (uint32_t v : particles) {
      V<T> pv = pos[v];
      T dmin = T(3.0e38);
      const auto& tr = vtTrack[v];
      for (int i = 0; i < 4; i++) {
        uint32_t t = tr[i];
        if (t >= nTris) break;
        const auto& tv = tris[t];
        if (in2(v, tv[0]) || in2(v, tv[1]) || in2(v, tv[2])) continue;
        V<T> bary;
        V<T> q = closestTri(pv, pos[tv[0]], pos[tv[1]], pos[tv[2]], bary);
        V<T> dd = pv - q;
        dmin = std::min(dmin, dotv(dd, dd));
      }
      uint32_t cnt = std::min<uint32_t>(vEdgeCount[v], 6u);
      for (uint32_t k = 0; k < cnt; k++) {
        uint32_t me = vEdgeList[vEdgeStart[v] + k];
        if (me >= numContactEdges) continue;
        uint32_t other = eeTrack[me][0];
        if (other >= numContactEdges) continue;
        const auto eo = edges[other];
        if (in2(v, eo.first) || in2(v, eo.second)) continue;
        const auto em_ = edges[me];
        T s_, t_;
        V<T> a0 = pos[em_.first], a1 = pos[em_.second], b0 = pos[eo.first], b1 = pos[eo.second];
        segSeg(a0, a1, b0, b1, s_, t_);
        V<T> dd = (a0 + (a1 - a0) * s_) - (b0 + (b1 - b0) * t_);
        dmin = std::min(dmin, dotv(dd, dd));
      }
      ogcBound[v] = dmin;
      ogcPrev[v] = pos[v];
    }
  }

  void datContacts() {
    for (const auto& pr : datPairs) {
      Soft r;
      if (!datContact(pr, r)) continue;
      softCandidates++;
      if (soft.size() < maxSoft) soft.push_back(r);
    }
  }

  // -------------------------------------------------------------------------
  // Planar-DAT truncation (27_planar_dat.metal:536-560 plane fraction,
  // 291-337 fallback normals, 756-906 pair reduction, 1176-1261 apply,
  // 1289-1310 failed-surface restore; order GPUSolver.swift:3824-3871 after
  // the predictor and 4196-4229 after every color).  Each pair places a
  // division plane between its two sides at the anchor pose; every vertex
  // approaching the plane keeps the fraction of its displacement that stops
  // it short of the plane (relaxation 0.85, time margin 1e-3), the minimum
  // over its pairs, and its displacement is capped at 0.5*0.85*rq.
  // -------------------------------------------------------------------------
  static constexpr double kDatGeomEps = 1.0e-8, kDatTimeMargin = 1.0e-3, kDatRelax = 0.85;
  T planeFraction(const V<T>& x, const V<T>& dx, const V<T>& n, const V<T>& plane, T side, bool report) {
    T s0 = side * dotv(n, x - plane);
    T q = side * dotv(n, dx);
    if (!std::isfinite(s0) || !std::isfinite(q)) {
      if (report) { datInvalid++; datNonfinite++; }
      return T(0);
    }
    if (s0 < -(T)kDatGeomEps) return T(0);
    if (q >= T(0) || s0 + q > T(0)) return T(1);
    T hit = std::max(s0, T(0)) / (-q);
    if (hit > T(1) + (T)kDatGeomEps) return T(1);
    return clampT(std::min((T)kDatRelax * hit, hit - (T)kDatTimeMargin), T(0), T(1));
  }
  void lowerT(uint32_t v, T t) {
    if (t < datT[v]) datT[v] = t;
  }
  void atomicT(uint32_t v, const V<T>& x, const V<T>& dx, const V<T>& n, const V<T>& plane, T side) {
    T t = planeFraction(x, dx, n, plane, side, true);
    if (t < T(1)) lowerT(v, t);
  }
  static bool vtFallbackNormal(const V<T>& a, const V<T>& b, const V<T>& c, const V<T>& rel, V<T>& n) {
    V<T> raw = crossv(b - a, c - a);
    T mag = lenv(raw);
    if (!(std::isfinite(mag) && mag >= (T)kDatGeomEps)) return false;
    n = raw / mag;
    if (dotv(n, rel) > T(0)) n = -n;
    return true;
  }
  static bool eeFallbackNormal(const V<T>& a0, const V<T>& a1, const V<T>& b0, const V<T>& b1, const V<T>& rel,
                               V<T>& n) {
    V<T> ua = a1 - a0, ub = b1 - b0;
    V<T> raw = crossv(ua, ub);
    T mag = lenv(raw);
    if (std::isfinite(mag) && mag >= (T)kDatGeomEps) {
      n = raw / mag;
    } else {
      T la = lenv(ua), lb = lenv(ub);
      V<T> u = la >= lb ? ua : ub;
      T ul = std::max(la, lb);
      if (!std::isfinite(ul) || ul < (T)kDatGeomEps) return false;
      u = u / ul;
      V<T> perp = rel - u * dotv(u, rel);
      T pl = lenv(perp);
      if (!std::isfinite(pl) || pl < (T)kDatGeomEps) {
        V<T> axis = std::fabs(u.x) < T(0.7) ? V<T>(1, 0, 0) : V<T>(0, 1, 0);
        perp = crossv(u, axis);
        pl = lenv(perp);
      }
      if (!std::isfinite(pl) || pl < (T)kDatGeomEps) return false;
      n = perp / pl;
    }
    if (dotv(n, rel) > T(0)) n = -n;
    return finite3(n);
  }
  // one pair's plane fractions (dat_reduce); color < 0 is the full pass
  void datReducePair(const std::array<uint32_t, 3>& pr, const std::vector<V<T>>& anchor, int color) {
    auto active = [&](uint32_t b) { return color < 0 || (mass[b] > T(0) && (int)colors[b] == color); };
    const T rq = (T)datQueryRadius, eps = (T)kDatGeomEps;
    if (pr[0] == 0) {
      uint32_t v = pr[1];
      const auto& t = tris[pr[2]];
      bool aV = active(v), aA = active(t[0]), aB = active(t[1]), aC = active(t[2]);
      if (!aV && !aA && !aB && !aC) return;
      V<T> xv = anchor[v], a = anchor[t[0]], b = anchor[t[1]], c = anchor[t[2]];
      V<T> bary;
      V<T> cp = closestTri(xv, a, b, c, bary);
      V<T> nh = xv - cp;
      T gap = lenv(nh);
      if (!std::isfinite(gap)) {
        datInvalid++;
        datNonfinite++;
        lowerT(v, T(0)); lowerT(t[0], T(0)); lowerT(t[1], T(0)); lowerT(t[2], T(0));
        return;
      }
      if (gap > rq) return;
      V<T> dv = pos[v] - xv, da = pos[t[0]] - a, db = pos[t[1]] - b, dc = pos[t[2]] - c;
      V<T> n;
      if (gap >= eps) {
        n = nh / gap;
      } else {
        if (!vtFallbackNormal(a, b, c, dv - (da + db + dc) / T(3), n)) {
          datInvalid++;
          datVTDeg++;
          lowerT(v, T(0)); lowerT(t[0], T(0)); lowerT(t[1], T(0)); lowerT(t[2], T(0));
          return;
        }
        datVTDeg++;
      }
      T apV = std::max(-dotv(n, dv), T(0));
      T apT = std::max(std::max(dotv(n, da), dotv(n, db)), std::max(dotv(n, dc), T(0)));
      T lambda = apV + apT == T(0) ? T(0.5) : clampT(apT / (apV + apT), T(0.05), T(0.95));
      V<T> plane = gap >= eps ? cp + nh * lambda : cp;
      if (apV > T(0) && aV) atomicT(v, xv, dv, n, plane, T(1));
      if (apT > T(0)) {
        if (aA) atomicT(t[0], a, da, n, plane, T(-1));
        if (aB) atomicT(t[1], b, db, n, plane, T(-1));
        if (aC) atomicT(t[2], c, dc, n, plane, T(-1));
      }
      return;
    }
    auto ea = edges[pr[1]], eb = edges[pr[2]];
    bool aA0 = active(ea.first), aA1 = active(ea.second), aB0 = active(eb.first), aB1 = active(eb.second);
    if (!aA0 && !aA1 && !aB0 && !aB1) return;
    V<T> a0 = anchor[ea.first], a1 = anchor[ea.second], b0 = anchor[eb.first], b1 = anchor[eb.second];
    T s_, t_;
    segSeg(a0, a1, b0, b1, s_, t_);
    V<T> ca = a0 + (a1 - a0) * s_, cb = b0 + (b1 - b0) * t_;
    V<T> nh = ca - cb;
    T gap = lenv(nh);
    if (!std::isfinite(gap)) {
      datInvalid++;
      datNonfinite++;
      lowerT(ea.first, T(0)); lowerT(ea.second, T(0)); lowerT(eb.first, T(0)); lowerT(eb.second, T(0));
      return;
    }
    if (gap > rq) return;
    V<T> da0 = pos[ea.first] - a0, da1 = pos[ea.second] - a1, db0 = pos[eb.first] - b0, db1 = pos[eb.second] - b1;
    V<T> n;
    if (gap > eps) {
      n = nh / gap;
    } else {
      if (!eeFallbackNormal(a0, a1, b0, b1, (da0 + da1 - db0 - db1) * T(0.5), n)) {
        datInvalid++;
        datEEDeg++;
        lowerT(ea.first, T(0)); lowerT(ea.second, T(0)); lowerT(eb.first, T(0)); lowerT(eb.second, T(0));
        return;
      }
      datEEDeg++;
    }
    T apA = std::max(std::max(-dotv(n, da0), -dotv(n, da1)), T(0));
    T apB = std::max(std::max(dotv(n, db0), dotv(n, db1)), T(0));
    T lambda = apA + apB == T(0) ? T(0.5) : clampT(apB / (apA + apB), T(0.05), T(0.95));
    V<T> plane = gap > eps ? cb + nh * lambda : cb;
    if (apA > T(0)) {
      if (aA0) atomicT(ea.first, a0, da0, n, plane, T(1));
      if (aA1) atomicT(ea.second, a1, da1, n, plane, T(1));
    }
    if (apB > T(0)) {
      if (aB0) atomicT(eb.first, b0, db0, n, plane, T(-1));
      if (aB1) atomicT(eb.second, b1, db1, n, plane, T(-1));
    }
  }
  // apply one vertex's fraction and the displacement cap (dat_apply / dat_apply_color)
  void datApply(uint32_t v, const std::vector<V<T>>& anchor) {
    V<T> d = pos[v] - anchor[v];
    bool unsafeCap = datTotal > maxDatPairs || datPeak > maxDatPairs || datOverflow > 0 || datInvalid > 0;
    T t = datT[v];
    datT[v] = T(1);
    if (unsafeCap) t = T(0);
    if (!finite3(d) || !std::isfinite(t)) {
      datInvalid++;
      datNonfinite++;
      d = V<T>();
      t = T(0);
    }
    bool truncated = t < T(1);
    d = d * clampT(t, T(0), T(1));
    T maxD = T(0.5) * (T)kDatRelax * (T)datQueryRadius;
    T d2 = dotv(d, d);
    if (d2 > maxD * maxD) {
      d = d * (maxD / std::sqrt(d2));
      truncated = true;
    }
    if (truncated) {
      pos[v] = anchor[v] + d;
      lo[v] = V<T>();
      datTrunc++;
    }
  }
  // the full pass after the predictor: every pair, every surface particle
  void datPredictorPass() {
    for (const auto& pr : datPairs) datReducePair(pr, initLin, -1);
    for (uint32_t v : particles) {
      datApply(v, initLin);
      snapLin[v] = pos[v];  // the first iteration starts from the truncated pose
    }
  }
  // the pass after color c's primal: the color's surface particles
  void datColorPass(int c) {
    for (const auto& pr : datPairs) datReducePair(pr, datAnchor, c);
    for (uint32_t v : buckets[c])
      if (clothGroup[v] != 0u) datApply(v, datAnchor);
  }

  // OGC contact row (40_solver.metal:1237-1300); `snap` reads same-color partners from the snapshot
  V<T> softC(const Soft& sc, uint32_t self, bool snap, V<T>& n, V<T>& t1, V<T>& t2, V<T>& rAW) const {
    bool rigidA = sc.flags & 1u, roundA = sc.flags & 32u;
    n = sc.n;
    orthoBasis(n, t1, t2);
    rAW = V<T>();
    V<T> C = V<T>(sc.C0, 0, 0) * (T(1) - (T)set.alpha);
    for (int k = 0; k < 4; k++) {
      uint32_t b = sc.ids[k];
      if (b == WORLD) continue;
      V<T> p = snap ? pL(b, self) : pos[b];
      V<T> dq = p - initLin[b];
      C = C + V<T>(dotv(n, dq), dotv(t1, dq), dotv(t2, dq)) * sc.w[k];
      if (k == 0 && rigidA) {
        Q<T> q = snap ? pA(b, self) : rot[b];
        rAW = roundA ? sc.anchor : qrot(q, sc.anchor);
        V<T> dqa = qsub(q, initAng[b]);
        C = C + V<T>(dotv(crossv(rAW, n), dqa), dotv(crossv(rAW, t1), dqa), dotv(crossv(rAW, t2), dqa)) * sc.w[k];
      }
    }
    return C;
  }
  V<T> softF(const Soft& sc, const V<T>& C, T& kN) const {
    V<T> F = mulv(sc.penalty, C) + sc.lambda;
    kN = sc.penalty.x;
    T rT = sc.anchor.x;
    uint32_t kind = (sc.flags >> 2) & 3u;
    if (rT > 0 && kind != 2u && C.x < T(-0.5) * rT) {
      T dsep = std::max(C.x + rT, T(0.02) * rT);
      T kp = T(0.25) * rT * rT * sc.penalty.x;
      F.x = -kp / dsep + sc.lambda.x;
      kN = kp / (dsep * dsep);
    }
    F.x = std::min(F.x, T(0));
    T bnd = std::fabs(F.x) * sc.friction, fs = std::sqrt(F.y * F.y + F.z * F.z);
    if (fs > bnd && fs > 0) { F.y = F.y * (bnd / fs); F.z = F.z * (bnd / fs); }
    return F;
  }
  void stampSoft(const Soft& sc, uint32_t self, Sys& s) const {
    V<T> n, t1, t2, rAW;
    V<T> C = softC(sc, self, true, n, t1, t2, rAW);
    T kN;
    V<T> F = softF(sc, C, kN);
    int slot = -1;
    for (int k = 0; k < 4; k++) if (sc.ids[k] == self) { slot = k; break; }
    if (slot < 0) return;
    T w = sc.w[slot];
    T K[3] = {kN, sc.penalty.y, sc.penalty.z};
    V<T> rows[3] = {n * w, t1 * w, t2 * w};
    for (int r = 0; r < 3; r++) {
      s.ll = s.ll + M<T>::outer(rows[r], rows[r]) * K[r];
      s.gl = s.gl + rows[r] * F[r];
      if (slot == 0 && (sc.flags & 1u)) {
        V<T> ja = crossv(rAW, rows[r]);
        s.aa = s.aa + M<T>::outer(ja, ja) * K[r];
        s.cr = s.cr + M<T>::outer(ja, rows[r]) * K[r];
        s.ga = s.ga + ja * F[r];
      }
    }
  }
  // StVK membrane (40_solver.metal:1088-1130)
  void stampMembrane(const Membrane& mb, uint32_t self, Sys& s) const {
    V<T> x0 = pL(mb.ids[0], self), x1 = pL(mb.ids[1], self), x2 = pL(mb.ids[2], self);
    T a = mb.dm[0], b = mb.dm[1], c = mb.dm[2], d = mb.dm[3];
    V<T> d0 = x1 - x0, d1 = x2 - x0;
    V<T> f1 = d0 * a + d1 * b, f2 = d0 * c + d1 * d;
    T E11 = T(0.5) * (dotv(f1, f1) - T(1)), E22 = T(0.5) * (dotv(f2, f2) - T(1)), E12 = T(0.5) * dotv(f1, f2);
    T trE = E11 + E22;
    T S11 = T(2) * mb.muA * E11 + mb.lamA * trE, S22 = T(2) * mb.muA * E22 + mb.lamA * trE, S12 = T(2) * mb.muA * E12;
    V<T> p1 = f1 * S11 + f2 * S12, p2 = f1 * S12 + f2 * S22;
    T wx, wy;
    if (self == mb.ids[1]) { wx = a; wy = c; }
    else if (self == mb.ids[2]) { wx = b; wy = d; }
    else { wx = -a - b; wy = -c - d; }
    s.gl = s.gl + p1 * wx + p2 * wy;
    V<T> g11 = f1 * wx, g22 = f2 * wy, g12 = (f2 * wx + f1 * wy) * T(0.5), gtr = g11 + g22;
    s.ll = s.ll + M<T>::outer(g11, g11) * (T(2) * mb.muA) + M<T>::outer(g22, g22) * (T(2) * mb.muA) +
           M<T>::outer(g12, g12) * (T(4) * mb.muA) + M<T>::outer(gtr, gtr) * mb.lamA;
  }
  // Bergou hinge (40_solver.metal:1136-1150), evaluated translation-free:
  // m = sum_{o != self} K_o (x_o - x_self), equal to sum_i K_i x_i since K is
  // orthogonal to the constants
  void stampBend(const Bend& bd, uint32_t self, Sys& s) const {
    V<T> xs = pos[self], m;
    T k = 0;
    for (int i = 0; i < 4; i++) {
      if (bd.ids[i] == self) { k = bd.K[i]; continue; }
      m = m + (pL(bd.ids[i], self) - xs) * bd.K[i];
    }
    s.gl = s.gl + m * (bd.s * k);
    s.ll = s.ll + M<T>::diag(V<T>(1, 1, 1) * (bd.s * k * k));
  }
  // element contact dual: quadratic stage, mass-scaled cap, penalties <= 1e6 (40_solver.metal:1689-1727)
  void dualSoft(Soft& sc) {
    V<T> n, t1, t2, rAW;
    V<T> C = softC(sc, 0, false, n, t1, t2, rAW);
    V<T> F = mulv(sc.penalty, C) + sc.lambda;
    F.x = std::min(F.x, T(0));
    T bnd = std::fabs(F.x) * sc.friction, fs = std::sqrt(F.y * F.y + F.z * F.z);
    if (fs > bnd && fs > 0) { F.y = F.y * (bnd / fs); F.z = F.z * (bnd / fs); }
    F.x = std::max(F.x, -sc.lamCap);
    sc.lambda = F;
    T bl = set.beta_lin;
    if (F.x < 0) sc.penalty.x = std::min(sc.penalty.x + bl * std::fabs(C.x), T(1e6));
    if (fs <= bnd) {
      sc.penalty.y = std::min(sc.penalty.y + bl * std::fabs(C.y), T(1e6));
      sc.penalty.z = std::min(sc.penalty.z + bl * std::fabs(C.z), T(1e6));
    }
  }

  // contact adjacency of this frame: per body, active manifold indices in order
  std::vector<std::vector<uint32_t>> contactAdj;

  Sys assemble(uint32_t body) const {
    Sys s{M<T>::zero(), M<T>::zero(), M<T>::zero(), V<T>(), V<T>()};
    T dt2 = (T)set.dt * (T)set.dt;
    T m = mass[body];
    s.ll = M<T>::diag(V<T>(1, 1, 1) * (m / dt2));
    M<T> In = worldInertia(rot[body], moment[body]) * (T(1) / dt2);
    s.aa = In;
    s.gl = ((pos[body] - inertLin[body]) + lo[body]) * (m / dt2);
    s.ga = In.mulv(qsub(rot[body], inertAng[body]));
    // reference order of the sorted CSR: joints, springs, manifolds, tets
    for (uint32_t e : staticAdj[body]) {
      uint32_t k = e >> 28, i = e & 0x0FFFFFFFu;
      if (k == FKJ && !joints[i].broken) stampJoint(joints[i], body, s);
      else if (k == FKS) stampSpring(springs[i], body, s);
    }
    for (uint32_t mi : contactAdj[body]) stampManifold(manifolds[mi], body, s);
    for (uint32_t e : staticAdj[body]) {
      uint32_t k = e >> 28, i = e & 0x0FFFFFFFu;
      if (k == FKT) stampTet(tets[i], body, s);
      else if (k == FKMEM) stampMembrane(membranes[i], body, s);
      else if (k == FKBEND) stampBend(bends[i], body, s);
    }
    if (!softAdj.empty())
      for (uint32_t si : softAdj[body]) stampSoft(soft[si], body, s);
    return s;
  }

  void primal(uint32_t body) {
    Sys s = assemble(body);
    bool part = radiusSigned[body] < 0;
    if (part) { s.aa = M<T>::diag(V<T>(1, 1, 1)); s.cr = M<T>::zero(); s.ga = V<T>(); }
    T A[6][6], b[6], x[6];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        A[i][j] = s.ll.a[i][j];
        A[3 + i][3 + j] = s.aa.a[i][j];
        A[3 + i][j] = s.cr.a[i][j];
        A[j][3 + i] = s.cr.a[i][j];
      }
    for (int i = 0; i < 3; i++) { b[i] = -s.gl[i]; b[3 + i] = -s.ga[i]; }
    solve6(A, b, x);
    V<T> dl(x[0], x[1], x[2]), da(x[3], x[4], x[5]);
    for (int i = 0; i < 6; i++) if (!std::isfinite(x[i])) return;
    T maxLin = T(0.35) * std::fabs(radiusSigned[body]);
    T l2 = dotv(dl, dl);
    if (l2 > maxLin * maxLin) dl = dl * (maxLin / std::sqrt(l2));
    T a2 = dotv(da, da);
    if (a2 > T(0.25)) da = da * (T(0.5) / std::sqrt(a2));
    if (part && isoOn) dl = ogcTruncate(body, pos[body], dl);
    // (pos, lo) + dl renormalised exactly (TwoSum per component)
    V<T> t = lo[body] + dl;
    for (int k = 0; k < 3; k++) {
      T a = pos[body][k], b = t[k];
      T sum = a + b, bb = sum - a;
      T err = (a - (sum - bb)) + (b - bb);
      if (k == 0) { pos[body].x = sum; lo[body].x = err; }
      else if (k == 1) { pos[body].y = sum; lo[body].y = err; }
      else { pos[body].z = sum; lo[body].z = err; }
    }
    if (!part) rot[body] = qaddw(rot[body], da);
  }

  // -------------------------------------------------------------------------
  // dual (40_solver.metal:1689-1910)
  // -------------------------------------------------------------------------
  void dual() {
    T alpha = set.alpha, bl = set.beta_lin, ba = set.beta_ang, lmax = set.lambda_max;
    for (auto& j : joints) {
      if (j.broken) continue;
      if (dotv(j.penaltyLin, j.penaltyLin) > 0) {
        V<T> pAw = j.a == WORLD ? j.rA : xformv(pos[j.a], rot[j.a], j.rA);
        V<T> C = pAw - xformv(pos[j.b], rot[j.b], j.rB);
        if (j.flags & 1u) {
          C = C - j.C0Lin * alpha;
          V<T> l = mulv(j.penaltyLin, C) + j.lambdaLin;
          j.lambdaLin = V<T>(clampT(l.x, -lmax, lmax), clampT(l.y, -lmax, lmax), clampT(l.z, -lmax, lmax));
        }
        T cap = std::min(j.stiffLin, T(1e10));
        for (int k = 0; k < 3; k++) j.penaltyLin[k] = std::min(j.penaltyLin[k] + std::fabs(C[k]) * bl, cap);
      }
      if (dotv(j.penaltyAng, j.penaltyAng) > 0) {
        Q<T> qA = j.a == WORLD ? Q<T>{0, 0, 0, 1} : rot[j.a];
        V<T> C;
        if (j.hinge) C = crossv(qrot(qmul(qA, j.restRel), j.axis), qrot(rot[j.b], j.axis)) * j.torqueArm;
        else C = qsub(qmul(qA, j.restRel), rot[j.b]) * j.torqueArm;
        if (j.flags & 2u) {
          C = C - j.C0Ang * alpha;
          V<T> l = mulv(j.penaltyAng, C) + j.lambdaAng;
          j.lambdaAng = V<T>(clampT(l.x, -lmax, lmax), clampT(l.y, -lmax, lmax), clampT(l.z, -lmax, lmax));
        }
        T cap = std::min(j.stiffAng, T(1e10));
        for (int k = 0; k < 3; k++) j.penaltyAng[k] = std::min(j.penaltyAng[k] + std::fabs(C[k]) * ba, cap);
      }
      T lin2 = (j.flags & 8u) ? dotv(j.lambdaLin, j.lambdaLin) : T(0);
      if ((j.flags & 4u) && dotv(j.lambdaAng, j.lambdaAng) + lin2 > j.fracture * j.fracture) {
        j.penaltyLin = j.penaltyAng = j.lambdaLin = j.lambdaAng = V<T>();
        j.broken = 1;
      }
    }
    for (auto& sp : springs) {
      if (!sp.hard) continue;
      T C = lenv(xformv(pos[sp.a], rot[sp.a], sp.rA) - xformv(pos[sp.b], rot[sp.b], sp.rB)) - sp.rest - sp.C0 * alpha;
      T big = std::numeric_limits<T>::max();
      T mm = std::min(mass[sp.a] > 0 ? mass[sp.a] : big, mass[sp.b] > 0 ? mass[sp.b] : big);
      T cap = std::min(lmax, std::max(T(2), T(5e3) * mm));
      sp.lambda = clampT(T(0.98) * sp.lambda + sp.penalty * C, T(0), cap);
      if (C > 0) sp.penalty = std::min(sp.penalty + C * bl, std::min(sp.stiffness, T(1e10)));
    }
    for (auto& m : manifolds) {
      if (m.n == 0) continue;
      bool rAr = (m.flags & 2u) != 0, rBr = (m.flags & 4u) != 0;
      V<T> dqAL = pos[m.a] - initLin[m.a], dqBL = pos[m.b] - initLin[m.b];
      V<T> dqAA = qsub(rot[m.a], initAng[m.a]), dqBA = qsub(rot[m.b], initAng[m.b]);
      T big = std::numeric_limits<T>::max();
      T mm = std::min(mass[m.a] > 0 ? mass[m.a] : big, mass[m.b] > 0 ? mass[m.b] : big);
      T cap = std::min(lmax, std::max(T(10), T(1e5) * mm));
      for (uint32_t i = 0; i < m.n; i++) {
        V<T> rAW = rAr ? m.c[i].rA : qrot(rot[m.a], m.c[i].rA);
        V<T> rBW = rBr ? m.c[i].rB : qrot(rot[m.b], m.c[i].rB);
        T fs, bnd;
        V<T> F;
        V<T> C = contactC(m, (int)i, dqAL, dqAA, dqBL, dqBA, rAW, rBW, fs, bnd, F);
        F.x = std::max(F.x, -cap);
        m.c[i].lambda = F;
        V<T> p = m.c[i].penalty;
        if (F.x < 0) p.x = std::min(p.x + bl * std::fabs(C.x), T(1e10));
        if (fs <= bnd) {
          p.y = std::min(p.y + bl * std::fabs(C.y), T(1e6));
          p.z = std::min(p.z + bl * std::fabs(C.z), T(1e6));
          m.c[i].stick = std::sqrt(C.y * C.y + C.z * C.z) < T(1e-5) ? T(1) : T(0);
        }
        m.c[i].penalty = p;
      }
    }
    for (auto& sc : soft) dualSoft(sc);
    snapLin = pos;
    snapAng = rot;
  }

  // -------------------------------------------------------------------------
  // dynamic coloring (40_solver.metal:320-540)
  // -------------------------------------------------------------------------
  void neighbors(uint32_t v, std::vector<uint32_t>& out) const {
    out.clear();
    for (uint32_t e : staticAdj[v]) {
      uint32_t k = e >> 28, i = e & 0x0FFFFFFFu;
      if (k == FKJ) out.push_back(joints[i].a == v ? joints[i].b : joints[i].a);
      else if (k == FKS) out.push_back(springs[i].a == v ? springs[i].b : springs[i].a);
      else if (k == FKT) for (int q = 0; q < 4; q++) if (tets[i].ids[q] != v) out.push_back(tets[i].ids[q]);
    }
    for (uint32_t mi : contactAdj[v]) out.push_back(manifolds[mi].a == v ? manifolds[mi].b : manifolds[mi].a);
  }
  void dynamicColors() {
    std::vector<uint32_t> src = colors, dst(nb, 0), nbrs;
    std::vector<uint8_t> changed(20, 0);
    for (int pass = 0; pass < 20; pass++) {
      if (pass > 0 && !changed[pass - 1]) { std::swap(src, dst); continue; }
      for (uint32_t v = 0; v < nb; v++) {
        if (!(mass[v] > 0)) { dst[v] = 0; continue; }
        uint32_t my = src[v];
        uint64_t lower = 0, all = 0;
        bool conflict = false;
        neighbors(v, nbrs);
        for (uint32_t n : nbrs) {
          if (!dyn(n)) continue;
          uint64_t bit = src[n] < 64 ? 1ull << src[n] : 0;
          all |= bit;
          if (n < v) { lower |= bit; if (src[n] == my) conflict = true; }
        }
        if (conflict) {
          uint64_t fr = ~lower;
          dst[v] = fr ? std::min<uint32_t>((uint32_t)__builtin_ctzll(fr), 63u) : 63u;
          changed[pass] = 1;
        } else {
          uint64_t fr = ~all;
          uint32_t best = fr ? (uint32_t)__builtin_ctzll(fr) : my;
          dst[v] = best < my ? best : my;
          if (best < my) changed[pass] = 1;
        }
      }
      std::swap(src, dst);
    }
    // src holds the final pass output (20 passes: back in the first buffer)
    if (changed[19]) {
      for (uint32_t v = 0; v < nb; v++) {
        if (!(mass[v] > 0)) { src[v] = 0; continue; }
        uint64_t lower = 0;
        neighbors(v, nbrs);
        for (uint32_t n : nbrs) if (dyn(n) && n < v && src[n] < 64) lower |= 1ull << src[n];
        uint64_t fr = ~lower;
        src[v] = fr ? (uint32_t)__builtin_ctzll(fr) : 63u;
      }
    }
    colors = src;
    colorConflicts = 0;
    for (uint32_t v = 0; v < nb; v++) {
      if (!(mass[v] > 0)) continue;
      neighbors(v, nbrs);
      for (uint32_t n : nbrs) if (dyn(n) && n < v && colors[n] == colors[v]) { colorConflicts++; break; }
    }
    rebuildBuckets();
  }

  // -------------------------------------------------------------------------
  // the step (GPUSolver.swift:3138-4405 order)
  // -------------------------------------------------------------------------
  // residual-driven frames (tolerance > 0): see stepToResidual
  double tolerance = 0, lastResidual = 0;
  int maxIterations = 200, checkEvery = 5, itersUsed = 0;
  // The bodies of one color are independent (same-color contact partners
  // come from the snapshot), so a color may be split over threads with
  // results identical to the serial order; threads = 1 is the serial sweep.
  int threads = 1;
  void sweep() {
    for (int c = 0; c < 64; c++) {
      const std::vector<uint32_t>& bk = buckets[c];
      const long n = (long)bk.size();
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1 && n > 256)
      for (long i = 0; i < n; i++) primal(bk[(size_t)i]);
      if (datOn && n > 0) datColorPass(c);
    }
    dual();
    if (isoOn) ogcRefresh();
  }
  void stepToResidual(double tol, int maxIt, int every) {
    tolerance = tol;
    maxIterations = maxIt;
    checkEvery = every;
    step();
    tolerance = 0;
  }

  void step() {
    // legacy servo rates advance their targets before the frame, wrapped to
    // [-pi, pi] in Float (GPUSolver.swift:4359-4416)
    for (auto& j : joints)
      if (j.rate != 0.0f) {
        const float twoPi = 6.28318530717958647692f;
        float t = (float)j.motorTarget + j.rate * (float)set.dt;
        t = t - twoPi * std::round(t / twoPi);
        j.motorTarget = (T)t;
      }
    // kinematic spinners turn before the frame: q <- normalize(dq q), dq the
    // rotation by omega dt about the spinner axis (GPUSolver.swift:4418-4427)
    for (size_t k = 0; k < spinBody.size(); k++) {
      float dq[4];
      oracle_spinner_rotation(spinAxisOmega[k].data(), set.dt, dq);
      Q<T>& q = rot[spinBody[k]];
      Q<T> r = qmul(Q<T>{(T)dq[0], (T)dq[1], (T)dq[2], (T)dq[3]}, q);
      T sc = T(1) / std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
      q = Q<T>{r.x * sc, r.y * sc, r.z * sc, r.w * sc};
    }
    frame++;
    broadphase();
    manifolds.assign(pairs.size(), Manifold<T>{});
    for (size_t i = 0; i < pairs.size(); i++) collide((uint32_t)i, manifolds[i]);
    soft.clear();
    softCandidates = softRT = 0;
    datOverflow = datNonfinite = 0;
    datPeak = datPeakVT = datInvalid = datTrunc = datVTDeg = datEEDeg = 0;
    // cloth: the pair query at the step-start pose, rigid-feature contacts,
    // the predictor, its full truncation pass, then the query at the accepted
    // predictor pose feeds the contacts and the per-color passes, anchored
    // there (GPUSolver.swift:3685-3870)
    if (nTris) {
      if (datOn) datBuild();
      rtEmit();
      if (isoOn) isoEmit();
    }
    warmstartConstraints();
    warmstartBodies();
    if (datOn) datPredictorPass();
    if (pairStream) {
      datBuild();
      datContacts();
    }
    if (datOn) for (uint32_t v : particles) datAnchor[v] = pos[v];
    softAdj.assign(nb, {});
    for (uint32_t i = 0; i < soft.size(); i++)
      for (int k = 0; k < 4; k++)
        if (dyn(soft[i].ids[k])) softAdj[soft[i].ids[k]].push_back(i);
    contactAdj.assign(nb, {});
    activeManifolds = 0;
    for (uint32_t i = 0; i < manifolds.size(); i++) {
      const Manifold<T>& m = manifolds[i];
      if (m.n == 0) continue;
      activeManifolds++;
      if (dyn(m.a)) contactAdj[m.a].push_back(i);
      if (dyn(m.b)) contactAdj[m.b].push_back(i);
    }
    if (dynamicColoring) dynamicColors();
    if (tolerance > 0) {
      // residual-driven: chunks of checkEvery sweeps until r <= tolerance
      itersUsed = 0;
      while (itersUsed < maxIterations) {
        int n = std::min(checkEvery, maxIterations - itersUsed);
        for (int it = 0; it < n; it++) sweep();
        itersUsed += n;
        lastResidual = residual();
        if (lastResidual <= tolerance) break;
      }
    } else {
      for (int it = 0; it < set.iterations; it++) sweep();
    }
    // a failed surface query restores the step-start surface
    // (dat_restore_failed_surface, 27_planar_dat.metal:1289-1310)
    if (datOn && (softCandidates > maxSoft || datPeak > maxDatPairs || datOverflow > 0 || datInvalid > 0))
      for (uint32_t i = 0; i < nb; i++)
        if (radiusSigned[i] < 0) {
          pos[i] = initLin[i];
          lo[i] = V<T>();
        }
    // BDF1 finalize (40_solver.metal:2108-2150)
    T dt = set.dt;
    T maxSpeed = std::max(T(30), T(1.5) * (T)cellSize / dt * T(0.5));
    for (uint32_t i = 0; i < nb; i++) {
      V<T> vOld = vel[i];
      prevVel[i] = vOld;
      if (!(mass[i] > 0)) continue;
      V<T> v = ((pos[i] - initLin[i]) + lo[i]) / dt;
      T s2 = dotv(v, v);
      if (s2 > maxSpeed * maxSpeed) v = v * (maxSpeed / std::sqrt(s2));
      if (radiusSigned[i] < 0 && set.particle_damping > 0) {
        V<T> bal = vOld + V<T>(0, 0, (T)set.gravity * gravityScale[i]) * dt;
        v = bal + (v - bal) / (T(1) + (T)set.particle_damping * dt);
      }
      if (radiusSigned[i] >= 0 && set.rigid_linear_damping > 0) v = v * std::exp(-(T)set.rigid_linear_damping * dt);
      V<T> om = qsub(rot[i], initAng[i]) / dt;
      if (radiusSigned[i] >= 0 && set.rigid_angular_damping > 0) om = om * std::exp(-(T)set.rigid_angular_damping * dt);
      vel[i] = v;
      angVel[i] = om;
    }
    // 1-ring velocity viscosity of surface particles (40_solver.metal:2154-2177)
    if (nTris && set.cloth_viscosity > 0) {
      std::vector<V<T>> v0 = vel;
      T wv = set.cloth_viscosity;
      for (uint32_t v : particles) {
        uint32_t n = nbrCount[v];
        if (n == 0) continue;
        V<T> avg;
        for (uint32_t k = nbrStart[v]; k < nbrStart[v] + n; k++) avg = avg + v0[nbrList[k]];
        avg = avg / (T)n;
        vel[v] = v0[v] + (avg - v0[v]) * wv;
      }
    }
    // persistence maps of this frame for the next one
    prevSoft = soft;
    softMap.clear();
    for (uint32_t i = 0; i < prevSoft.size(); i++) softMap.emplace(std::make_pair(prevSoft[i].keyA, prevSoft[i].keyB), (int)i);
    prevManifolds = manifolds;
    prevMap.clear();
    for (uint32_t i = 0; i < prevManifolds.size(); i++)
      if (prevManifolds[i].n) prevMap[{prevManifolds[i].ca, prevManifolds[i].cb}] = (int)i;
  }

  double residual() const {
    double g2 = 0, w2 = 0;
    // evaluated with live poses only (no same-color snapshot)
    Oracle<T>* self = const_cast<Oracle<T>*>(this);
    std::vector<V<T>> sl = snapLin;
    std::vector<Q<T>> sa = snapAng;
    self->snapLin = pos;
    self->snapAng = rot;
    for (uint32_t b = 0; b < nb; b++) {
      if (!(mass[b] > 0)) continue;
      Sys s = assemble(b);
      g2 += (double)dotv(s.gl, s.gl);
      if (radiusSigned[b] >= 0) g2 += (double)dotv(s.ga, s.ga);
      double mg = (double)mass[b] * (double)set.gravity * (double)gravityScale[b];
      w2 += mg * mg;
    }
    self->snapLin = sl;
    self->snapAng = sa;
    return w2 > 0 ? std::sqrt(g2 / w2) : std::sqrt(g2);
  }

  void setStates(const uint32_t* idx, uint32_t n, const float* p, const float* r, const float* v, const float* w) {
    std::vector<uint8_t> flag(nb, 0);
    for (uint32_t k = 0; k < n; k++) {
      uint32_t i = idx ? idx[k] : k;
      pos[i] = V<T>(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
      lo[i] = V<T>();
      rot[i] = Q<T>{r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]};
      vel[i] = V<T>(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
      angVel[i] = V<T>(w[3 * k], w[3 * k + 1], w[3 * k + 2]);
      prevVel[i] = vel[i];
      flag[i] = 1;
    }
    auto f = [&](uint32_t b) { return b != WORLD && flag[b]; };
    for (auto& j : joints)
      if (f(j.a) || f(j.b)) {
        j.lambdaLin = j.lambdaAng = V<T>();
        j.motorZ = 0; j.predTwist = 0; j.effort = 0;
        j.penaltyLin = j.penLin0;
        j.penaltyAng = j.penAng0;
      }
    for (auto& s : springs) if (f(s.a) || f(s.b)) { s.lambda = 0; s.penalty = 0; s.C0 = 0; }
    // incident manifolds (both frame buffers) become cold
    for (auto& m : prevManifolds) if (f(m.a) || f(m.b)) m = Manifold<T>{};
    for (auto& m : manifolds) if (f(m.a) || f(m.b)) m = Manifold<T>{};
  }
};

struct oracle_solver {
  int precision;
  Oracle<float>* f = nullptr;
  Oracle<double>* d = nullptr;
};

extern "C" {

int oracle_create(const avbd_scene_desc* scene, int precision, int maxPairsPerBody, oracle_solver** out) {
  if (!scene || !out) return 1;
  oracle_solver* o = new oracle_solver();
  o->precision = precision == 64 ? 64 : 32;
  int rc;
  if (o->precision == 64) { o->d = new Oracle<double>(); rc = o->d->init(scene, maxPairsPerBody >= 0 ? maxPairsPerBody : 16); }
  else { o->f = new Oracle<float>(); rc = o->f->init(scene, maxPairsPerBody >= 0 ? maxPairsPerBody : 16); }
  if (rc) { oracle_destroy(o); *out = nullptr; return rc; }
  *out = o;
  return 0;
}

void oracle_destroy(oracle_solver* o) {
  if (!o) return;
  delete o->f;
  delete o->d;
  delete o;
}

int oracle_step(oracle_solver* o) {
  if (o->f) o->f->step(); else o->d->step();
  return 0;
}

int oracle_step_to_residual(oracle_solver* o, double tol, int max_iterations, int check_every, int* iterations,
                            double* residual) {
  auto go = [&](auto* s) {
    s->stepToResidual(tol, max_iterations, check_every);
    if (iterations) *iterations = s->itersUsed;
    if (residual) *residual = s->lastResidual;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_set_joint_world_anchors(oracle_solver* o, const uint32_t* joints, uint32_t n, const float* pts) {
  // setJointWorldAnchors (GPUSolver.swift:2481-2489): the world point of a
  // world-anchored joint
  auto go = [&](auto* s) {
    for (uint32_t k = 0; k < n; k++) {
      if (joints[k] >= s->joints.size()) continue;
      s->joints[joints[k]].rA = decltype(s->joints[0].rA)(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_set_dat_test_hooks(oracle_solver* o, float queryRadius, int32_t pairCapacity, int32_t softCapacity) {
  auto go = [&](auto* s) {
    if (queryRadius > 0) s->datQueryRadius = queryRadius;
    if (pairCapacity > 0) s->maxDatPairs = std::min<uint32_t>(s->maxDatPairs, (uint32_t)pairCapacity);
    if (softCapacity > 0) s->maxSoft = std::min<uint32_t>(s->maxSoft, (uint32_t)softCapacity);
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_set_threads(oracle_solver* o, int threads) {
  if (threads < 1) threads = 1;
  if (o->f) o->f->threads = threads; else o->d->threads = threads;
  return 0;
}
int oracle_set_settings(oracle_solver* o, const avbd_settings* s) {
  if (o->f) o->f->set = *s; else o->d->set = *s;
  return 0;
}

}  // extern "C"

template <typename T>
static void getStates(Oracle<T>* s, uint32_t n, float* p, float* r, float* v, float* w) {
  for (uint32_t i = 0; i < n && i < s->nb; i++) {
    if (p) { p[3 * i] = (float)s->pos[i].x; p[3 * i + 1] = (float)s->pos[i].y; p[3 * i + 2] = (float)s->pos[i].z; }
    if (r) { r[4 * i] = (float)s->rot[i].x; r[4 * i + 1] = (float)s->rot[i].y; r[4 * i + 2] = (float)s->rot[i].z; r[4 * i + 3] = (float)s->rot[i].w; }
    if (v) { v[3 * i] = (float)s->vel[i].x; v[3 * i + 1] = (float)s->vel[i].y; v[3 * i + 2] = (float)s->vel[i].z; }
    if (w) { w[3 * i] = (float)s->angVel[i].x; w[3 * i + 1] = (float)s->angVel[i].y; w[3 * i + 2] = (float)s->angVel[i].z; }
  }
}
template <typename T>
static void getStates64(Oracle<T>* s, uint32_t n, double* p, double* r, double* v, double* w) {
  for (uint32_t i = 0; i < n && i < s->nb; i++) {
    if (p) { p[3 * i] = s->pos[i].x; p[3 * i + 1] = s->pos[i].y; p[3 * i + 2] = s->pos[i].z; }
    if (r) { r[4 * i] = s->rot[i].x; r[4 * i + 1] = s->rot[i].y; r[4 * i + 2] = s->rot[i].z; r[4 * i + 3] = s->rot[i].w; }
    if (v) { v[3 * i] = s->vel[i].x; v[3 * i + 1] = s->vel[i].y; v[3 * i + 2] = s->vel[i].z; }
    if (w) { w[3 * i] = s->angVel[i].x; w[3 * i + 1] = s->angVel[i].y; w[3 * i + 2] = s->angVel[i].z; }
  }
}

extern "C" {

int oracle_get_body_states(oracle_solver* o, uint32_t n, float* p, float* r, float* v, float* w) {
  if (o->f) getStates(o->f, n, p, r, v, w); else getStates(o->d, n, p, r, v, w);
  return 0;
}
int oracle_get_body_states_f64(oracle_solver* o, uint32_t n, double* p, double* r, double* v, double* w) {
  if (o->f) getStates64(o->f, n, p, r, v, w); else getStates64(o->d, n, p, r, v, w);
  return 0;
}
int oracle_set_body_states(oracle_solver* o, const uint32_t* idx, uint32_t n, const float* p, const float* r,
                           const float* v, const float* w) {
  if (o->f) o->f->setStates(idx, n, p, r, v, w); else o->d->setStates(idx, n, p, r, v, w);
  return 0;
}
int oracle_get_colors(oracle_solver* o, uint32_t n, uint32_t* c) {
  const std::vector<uint32_t>& col = o->f ? o->f->colors : o->d->colors;
  for (uint32_t i = 0; i < n && i < col.size(); i++) c[i] = col[i];
  return 0;
}
int oracle_get_cell_lists(oracle_solver* o, uint32_t* hs, uint32_t* nh, float* cs, uint32_t* start, uint32_t* count,
                          uint32_t* bodies) {
  auto go = [&](auto* s) {
    if (hs) *hs = s->hashSize;
    if (nh) *nh = (uint32_t)s->hashed.size();
    if (cs) *cs = s->cellSize;
    if (start) for (size_t i = 0; i < s->cellStart.size(); i++) start[i] = s->cellStart[i];
    if (count) for (size_t i = 0; i < s->cellCount.size(); i++) count[i] = s->cellCount[i];
    if (bodies) for (size_t i = 0; i < s->cellBodies.size(); i++) bodies[i] = s->cellBodies[i];
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_get_pairs(oracle_solver* o, uint32_t cap, uint32_t* n, uint32_t* pairs) {
  auto go = [&](auto* s) {
    *n = (uint32_t)s->pairs.size();
    for (uint32_t i = 0; i < *n && i < cap; i++) { pairs[2 * i] = s->pairs[i].first; pairs[2 * i + 1] = s->pairs[i].second; }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_get_manifolds(oracle_solver* o, uint32_t cap, uint32_t* n, uint32_t* h3, float* n3, float* lp) {
  auto go = [&](auto* s) {
    const auto& ms = s->prevManifolds;
    *n = (uint32_t)ms.size();
    for (uint32_t i = 0; i < *n && i < cap; i++) {
      if (h3) { h3[3 * i] = ms[i].a; h3[3 * i + 1] = ms[i].b; h3[3 * i + 2] = ms[i].n; }
      if (n3) { n3[3 * i] = (float)ms[i].nrm.x; n3[3 * i + 1] = (float)ms[i].nrm.y; n3[3 * i + 2] = (float)ms[i].nrm.z; }
      if (lp) { lp[2 * i] = ms[i].n ? (float)ms[i].c[0].lambda.x : 0.0f; lp[2 * i + 1] = ms[i].n ? (float)ms[i].c[0].penalty.x : 0.0f; }
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
// the last frame's manifolds as 704-byte records: header (a, b, count,
// flags), collider pair, (normal, dynamic friction), (t1, static friction),
// then 8 contacts of (rA, feature bits), (rB, stick), C0, lambda, penalty
int oracle_get_manifold_records(oracle_solver* o, uint32_t cap, uint32_t* n, void* records) {
  auto go = [&](auto* s) {
    const auto& ms = s->prevManifolds;
    *n = (uint32_t)ms.size();
    for (uint32_t i = 0; i < *n && i < cap && records; i++) {
      uint32_t* u = (uint32_t*)records + 176 * i;
      float* f = (float*)u;
      const auto& m = ms[i];
      u[0] = m.a; u[1] = m.b; u[2] = m.n; u[3] = m.flags;
      u[4] = m.ca; u[5] = m.cb; u[6] = 0; u[7] = 0;
      f[8] = (float)m.nrm.x; f[9] = (float)m.nrm.y; f[10] = (float)m.nrm.z; f[11] = (float)m.dynF;
      f[12] = (float)m.t1.x; f[13] = (float)m.t1.y; f[14] = (float)m.t1.z; f[15] = (float)m.statF;
      for (uint32_t k = 0; k < 8; k++) {
        float* c = f + 16 + 20 * k;
        for (int j = 0; j < 20; j++) c[j] = 0.0f;
        if (k >= m.n) continue;
        const auto& q = m.c[k];
        c[0] = (float)q.rA.x; c[1] = (float)q.rA.y; c[2] = (float)q.rA.z; memcpy(&c[3], &q.feature, 4);
        c[4] = (float)q.rB.x; c[5] = (float)q.rB.y; c[6] = (float)q.rB.z; c[7] = (float)q.stick;
        c[8] = (float)q.C0.x; c[9] = (float)q.C0.y; c[10] = (float)q.C0.z;
        c[12] = (float)q.lambda.x; c[13] = (float)q.lambda.y; c[14] = (float)q.lambda.z;
        c[16] = (float)q.penalty.x; c[17] = (float)q.penalty.y; c[18] = (float)q.penalty.z;
      }
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
// feature keys of the last frame's contacts, 8 per manifold (0xFFFFFFFF past the count)
int oracle_get_manifold_keys(oracle_solver* o, uint32_t cap, uint32_t* n, uint32_t* keys8) {
  auto go = [&](auto* s) {
    const auto& ms = s->prevManifolds;
    *n = (uint32_t)ms.size();
    for (uint32_t i = 0; i < *n && i < cap && keys8; i++)
      for (uint32_t k = 0; k < 8; k++) keys8[8 * i + k] = k < ms[i].n ? ms[i].c[k].feature : 0xFFFFFFFFu;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_get_iso_state(oracle_solver* o, uint32_t* vt_track, uint32_t* ee_track, float* bound, float* anchor,
                         uint32_t* counts) {
  auto go = [&](auto* s) {
    if (!s->isoOn) return -1;
    for (uint32_t i = 0; i < s->nb; i++) {
      for (int k = 0; k < 4 && vt_track; k++) vt_track[4 * i + k] = s->vtTrack[i][k];
      if (bound) bound[i] = (float)s->ogcBound[i];
      for (int k = 0; k < 3 && anchor; k++) anchor[3 * i + k] = (float)s->ogcPrev[i][k];
    }
    for (uint32_t e = 0; e < s->numContactEdges && ee_track; e++)
      for (int k = 0; k < 4; k++) ee_track[4 * e + k] = s->eeTrack[e][k];
    if (counts) {
      counts[0] = s->ogcTruncFrame;
      counts[1] = s->ogcRefreshFrame;
    }
    return 0;
  };
  return o->f ? go(o->f) : go(o->d);
}
int oracle_get_motor_targets(oracle_solver* o, uint32_t n, float* targets) {
  auto go = [&](auto* s) {
    for (uint32_t i = 0; i < n && i < s->joints.size(); i++) targets[i] = (float)s->joints[i].motorTarget;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}

int oracle_get_collision_tris(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* tris) {
  auto go = [&](auto* s) {
    *n = (uint32_t)s->tris.size();
    for (uint32_t i = 0; i < std::min(*n, capacity); i++)
      for (int k = 0; k < 3; k++) tris[3 * i + k] = s->tris[i][k];
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}

int oracle_get_cloth_info(oracle_solver* o, avbd_cloth_info* info, uint32_t* edges, uint32_t* particles) {
  auto go = [&](auto* s) {
    if (info) {
      info->num_tris = s->nTris;
      info->num_edges = (uint32_t)s->edges.size();
      info->num_particles = (uint32_t)s->particles.size();
      info->num_membranes = (uint32_t)s->membranes.size();
      info->num_bends = (uint32_t)s->bends.size();
      info->elem_hash_size = 0;
      info->max_soft = s->maxSoft;
      info->soft_map_capacity = 0;
      info->max_dat_pairs = s->maxDatPairs;
      info->elem_cell_size = s->elemCellSize;
      info->elem_margin = s->elemMargin;
      info->dat_query_radius = s->datQueryRadius;
      info->surface_contact_cell_size = s->surfaceContactCellSize;
      info->num_contact_edges = s->numContactEdges;
      info->pair_stream = s->pairStream ? 1u : 0u;
      info->dat_on = s->datOn ? 1u : 0u;
      info->iso_on = s->isoOn ? 1u : 0u;
    }
    if (edges)
      for (size_t i = 0; i < s->edges.size(); i++) { edges[2 * i] = s->edges[i].first; edges[2 * i + 1] = s->edges[i].second; }
    if (particles)
      for (size_t i = 0; i < s->particles.size(); i++) particles[i] = s->particles[i];
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_get_dat_pairs(oracle_solver* o, uint32_t cap, uint32_t* n, uint32_t* pairs) {
  auto go = [&](auto* s) {
    *n = (uint32_t)s->datPairs.size();
    for (uint32_t i = 0; i < *n && i < cap && pairs; i++)
      for (int k = 0; k < 3; k++) pairs[3 * i + k] = s->datPairs[i][k];
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
// records in the backend's 112-byte SoftContactGPU layout (floats, ids and keys as bits)
int oracle_get_soft_contacts(oracle_solver* o, uint32_t cap, uint32_t* n, void* recs) {
  auto go = [&](auto* s) {
    *n = (uint32_t)s->prevSoft.size();
    if (!recs) return;
    for (uint32_t i = 0; i < *n && i < cap; i++) {
      const auto& r = s->prevSoft[i];
      uint32_t* u = (uint32_t*)recs + 28 * i;
      float* f = (float*)u;
      for (int k = 0; k < 4; k++) u[k] = r.ids[k];
      f[4] = (float)r.n.x; f[5] = (float)r.n.y; f[6] = (float)r.n.z; f[7] = (float)r.friction;
      f[8] = (float)r.anchor.x; f[9] = (float)r.anchor.y; f[10] = (float)r.anchor.z;
      memcpy(&f[11], &r.flags, 4);
      for (int k = 0; k < 4; k++) f[12 + k] = (float)r.w[k];
      f[16] = (float)r.C0; f[17] = 0.0f; f[18] = 0.0f; f[19] = (float)r.lamCap;
      f[20] = (float)r.lambda.x; f[21] = (float)r.lambda.y; f[22] = (float)r.lambda.z;
      memcpy(&f[23], &r.keyA, 4);
      f[24] = (float)r.penalty.x; f[25] = (float)r.penalty.y; f[26] = (float)r.penalty.z;
      memcpy(&f[27], &r.keyB, 4);
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
// the pair stream the element grid yields for the given positions (the
// solver state is left unchanged)
int oracle_dat_pairs_at(oracle_solver* o, const float* pos3, uint32_t cap, uint32_t* n, uint32_t* pairs) {
  auto go = [&](auto* s) {
    auto saved = s->pos;
    auto savedPairs = s->datPairs;
    for (uint32_t i = 0; i < s->nb; i++) {
      s->pos[i].x = pos3[3 * i];
      s->pos[i].y = pos3[3 * i + 1];
      s->pos[i].z = pos3[3 * i + 2];
    }
    s->datBuild();
    *n = (uint32_t)s->datPairs.size();
    for (uint32_t i = 0; i < *n && i < cap && pairs; i++)
      for (int k = 0; k < 3; k++) pairs[3 * i + k] = s->datPairs[i][k];
    s->pos = saved;
    s->datPairs = savedPairs;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_cloth_stats(oracle_solver* o, int32_t* datPairs, int32_t* datVT, int32_t* softCandidates, int32_t* softRT,
                       int32_t* overflow, int32_t* nonfinite) {
  auto go = [&](auto* s) {
    if (datPairs) *datPairs = (int32_t)s->datPeak;
    if (datVT) *datVT = (int32_t)s->datPeakVT;
    if (softCandidates) *softCandidates = (int32_t)s->softCandidates;
    if (softRT) *softRT = (int32_t)s->softRT;
    if (overflow) *overflow = (int32_t)s->datOverflow;
    if (nonfinite) *nonfinite = (int32_t)s->datNonfinite;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_dat_stats(oracle_solver* o, int32_t* out) {
  auto go = [&](auto* s) {
    out[0] = (int32_t)s->datTrunc;
    out[1] = (int32_t)s->datInvalid;
    out[2] = (int32_t)s->datVTDeg;
    out[3] = (int32_t)s->datEEDeg;
    out[4] = (int32_t)s->datNonfinite;
    out[5] = (int32_t)s->datOverflow;
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_residual(oracle_solver* o, double* out) {
  *out = o->f ? o->f->residual() : o->d->residual();
  return 0;
}
int oracle_stats(oracle_solver* o, int32_t* np, int32_t* am, int32_t* nc, int32_t* pc, int32_t* cc) {
  auto go = [&](auto* s) {
    if (np) *np = (int32_t)s->pairs.size();
    if (am) *am = (int32_t)s->activeManifolds;
    if (pc) *pc = (int32_t)s->pairCandidates;
    if (cc) *cc = (int32_t)s->colorConflicts;
    if (nc) {
      int used = 0;
      for (int c = 0; c < 64; c++) if (!s->buckets[c].empty()) used = c + 1;
      *nc = used;
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}
int oracle_map_find(oracle_solver* o, const uint32_t* pairs, uint32_t n, int32_t* out) {
  auto go = [&](auto* s) {
    for (uint32_t i = 0; i < n; i++) {
      auto it = s->prevMap.find({pairs[2 * i], pairs[2 * i + 1]});
      out[i] = it == s->prevMap.end() ? -1 : it->second;
    }
  };
  if (o->f) go(o->f); else go(o->d);
  return 0;
}

// Random.swift:3-21
void oracle_splitmix64(uint64_t seed, uint32_t n, uint64_t* out, float* outf) {
  uint64_t s = seed;
  for (uint32_t i = 0; i < n; i++) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    if (out) out[i] = z;
    if (outf) outf[i] = (float)(z >> 40) / (float)(1u << 24);
  }
}

// 40_solver.metal:965-1020 in double, plus the SNH energy for gradient checks
void oracle_snh_tet(const double* x, const double* dm, double mu, double lam, double vol, int self, double* energy,
                    double* grad3, double* hess9) {
  V<double> X[4];
  for (int i = 0; i < 4; i++) X[i] = V<double>(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
  V<double> r[3] = {V<double>(dm[0], dm[1], dm[2]), V<double>(dm[3], dm[4], dm[5]), V<double>(dm[6], dm[7], dm[8])};
  double muV = mu * vol, lamV = lam * vol, aT = 1.0 + muV / std::max(lamV, 1e-9);
  V<double> d0 = X[1] - X[0], d1 = X[2] - X[0], d2 = X[3] - X[0];
  V<double> f[3];
  for (int c = 0; c < 3; c++) f[c] = d0 * r[0][c] + d1 * r[1][c] + d2 * r[2][c];
  V<double> j0 = crossv(f[1], f[2]), j1 = crossv(f[2], f[0]), j2 = crossv(f[0], f[1]);
  double J = dotv(f[0], j0);
  double Ic = dotv(f[0], f[0]) + dotv(f[1], f[1]) + dotv(f[2], f[2]);
  if (energy) *energy = 0.5 * muV * (Ic - 3.0) + 0.5 * lamV * (J - aT) * (J - aT);
  double s2 = lamV * (J - aT);
  V<double> p0 = f[0] * muV + j0 * s2, p1 = f[1] * muV + j1 * s2, p2 = f[2] * muV + j2 * s2;
  V<double> w = self == 1 ? r[0] : self == 2 ? r[1] : self == 3 ? r[2] : -(r[0] + r[1] + r[2]);
  V<double> g = p0 * w.x + p1 * w.y + p2 * w.z;
  V<double> gj = j0 * w.x + j1 * w.y + j2 * w.z;
  if (grad3) { grad3[0] = g.x; grad3[1] = g.y; grad3[2] = g.z; }
  if (hess9)
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) hess9[3 * i + k] = (i == k ? muV * dotv(w, w) : 0.0) + lamV * gj[i] * gj[k];
}

void oracle_solve6x6(const double* A36, const double* b6, double* x6, int precision) {
  if (precision == 64) {
    double A[6][6], b[6];
    for (int i = 0; i < 36; i++) A[i / 6][i % 6] = A36[i];
    for (int i = 0; i < 6; i++) b[i] = b6[i];
    solve6(A, b, x6);
  } else {
    float A[6][6], b[6], x[6];
    for (int i = 0; i < 36; i++) A[i / 6][i % 6] = (float)A36[i];
    for (int i = 0; i < 6; i++) b[i] = (float)b6[i];
    solve6(A, b, x);
    for (int i = 0; i < 6; i++) x6[i] = x[i];
  }
}

// StVK membrane in double (40_solver.metal:1088-1130): positions x[9]
// (vertices 0..2), inverse rest shape dm[4] column major, mu A, lambda A,
// self slot; energy mu A |E|^2 + lambda A / 2 tr(E)^2, gradient and the
// Gauss-Newton block on self
void oracle_membrane(const double* x, const double* dm, double muA, double lamA, int self, double* energy,
                     double* grad3, double* hess9) {
  V<double> x0(x[0], x[1], x[2]), x1(x[3], x[4], x[5]), x2(x[6], x[7], x[8]);
  double a = dm[0], b = dm[1], c = dm[2], d = dm[3];
  V<double> d0 = x1 - x0, d1 = x2 - x0, f1 = d0 * a + d1 * b, f2 = d0 * c + d1 * d;
  double E11 = 0.5 * (dotv(f1, f1) - 1), E22 = 0.5 * (dotv(f2, f2) - 1), E12 = 0.5 * dotv(f1, f2), tr = E11 + E22;
  if (energy) *energy = muA * (E11 * E11 + E22 * E22 + 2 * E12 * E12) + 0.5 * lamA * tr * tr;
  double S11 = 2 * muA * E11 + lamA * tr, S22 = 2 * muA * E22 + lamA * tr, S12 = 2 * muA * E12;
  V<double> p1 = f1 * S11 + f2 * S12, p2 = f1 * S12 + f2 * S22;
  double wx = self == 1 ? a : self == 2 ? b : -a - b, wy = self == 1 ? c : self == 2 ? d : -c - d;
  V<double> g = p1 * wx + p2 * wy;
  if (grad3) { grad3[0] = g.x; grad3[1] = g.y; grad3[2] = g.z; }
  V<double> g11 = f1 * wx, g22 = f2 * wy, g12 = (f2 * wx + f1 * wy) * 0.5, gt = g11 + g22;
  M<double> H = M<double>::outer(g11, g11) * (2 * muA) + M<double>::outer(g22, g22) * (2 * muA) +
                M<double>::outer(g12, g12) * (4 * muA) + M<double>::outer(gt, gt) * lamA;
  if (hess9) for (int i = 0; i < 9; i++) hess9[i] = H.a[i / 3][i % 3];
}

// Hinge constants exactly as the upload derives them: a two-triangle scene
// (x[12] = edge ends v0, v1 then the opposite vertices), bends k1, k2;
// returns K[4] (scaled to the cotangent convention) and the stiffness s,
// or 0 when the upload would skip the hinge
int oracle_hinge(const float* x, float k1, float k2, float* K4, float* s_out) {
  avbd_scene_desc d;
  memset(&d, 0, sizeof(d));
  d.abi_version = AVBD_ABI_VERSION;
  d.settings.dt = 1.0f / 60.0f;
  float size[12], dens[4] = {1, 1, 1, 1}, fr[4] = {0.5f, 0.5f, 0.5f, 0.5f};
  uint32_t shape[4] = {AVBD_SHAPE_SPHERE, AVBD_SHAPE_SPHERE, AVBD_SHAPE_SPHERE, AVBD_SHAPE_SPHERE};
  uint8_t part[4] = {1, 1, 1, 1};
  for (int i = 0; i < 12; i++) size[i] = 0.002f;
  uint32_t tri[6] = {0, 1, 2, 0, 1, 3};
  float mu[2] = {1, 1}, bend[2] = {k1, k2};
  d.num_bodies = 4; d.body_size = size; d.body_density = dens; d.body_friction = fr; d.body_position = x;
  d.body_shape = shape; d.body_is_particle = part;
  d.num_tris = 2; d.tri_ids = tri; d.tri_mu = mu; d.tri_lambda = mu; d.tri_bend = bend;
  Oracle<float> o;
  o.set = d.settings;
  o.nb = 4;
  o.particle.assign(4, 1);
  o.clothGroup.assign(4, 0);
  o.clothInit(&d);
  if (o.bends.empty()) return 0;
  for (int k = 0; k < 4; k++) K4[k] = o.bends[0].K[k];
  *s_out = o.bends[0].s;
  return 1;
}

uint32_t oracle_cell_hash(float px, float py, float pz, float cs, uint32_t group, uint32_t hs) {
  return cellHash32((int)std::floor(px / cs), (int)std::floor(py / cs), (int)std::floor(pz / cs), group, hs);
}

}  // extern "C"
