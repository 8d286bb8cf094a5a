// cpu_path.cpp -- TEST INFRASTRUCTURE ONLY.  The reference's own CPU AVBD
// path restated (Sources/PhysicsAVBD/CPU/CPUSolver.swift, CPUManifold.swift,
// CPUForces.swift, the box / sphere branches of CPUCollide.swift, and
// Sources/SimCore/Math/Math.swift): the semantics of `scene.makeCPUSolver()`
// (PhysicsScene+CPUSolver.swift:4-61), which differ from the GPU kernels at
// these forks:
//   * plain LDL^T of the 6x6 block system (Math.swift:163-226), no
//     preconditioning or pivot clamp, no trust region;
//   * O(n^2) body-sphere broad phase over bounding radii, a manifold per
//     overlapping pair not already joined by a force (CPUSolver.swift:218-228);
//   * index-order Gauss-Seidel over the bodies, forces in creation order;
//   * box-box SAT without a speculative margin, spheres with the fixed margin;
//   * contact dual capped at lambdaMax (CPUManifold.swift:193-196);
//   * joints start at the penalty floor, hinge angular block diagonal
//     (CPUForces.swift:121-131); no motors or limits;
//   * springs bilateral (CPUForces.swift:212-242); no speed clamp.
// Bodies collide by their own shape (box or sphere).  Nothing in the product
// links this file; it is the CPU baseline of the stack workload and the
// cross-check of the reference's CPU-vs-GPU test.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <limits>
#include <vector>

#include "oracle.h"

namespace {

template <typename T>
struct V3 {
  T x = 0, y = 0, z = 0;
  V3() = default;
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> V3<T> mulc(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> T len(V3<T> a) { return std::sqrt(dot(a, a)); }
template <typename T> V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }

// quaternion (imag xyz, real w), Hamilton product
template <typename T>
struct Q {
  T x = 0, y = 0, z = 0, w = 1;
};
template <typename T> Q<T> qmul(Q<T> a, Q<T> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename T> Q<T> qnorm(Q<T> q) {
  T n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return {q.x / n, q.y / n, q.z / n, q.w / n};
}
template <typename T> Q<T> qconj(Q<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <typename T> Q<T> qinv(Q<T> q) {
  T n = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  return {-q.x / n, -q.y / n, -q.z / n, q.w / n};
}
template <typename T> V3<T> rot(Q<T> q, V3<T> v) {
  Q<T> r = qmul(qmul(q, Q<T>{v.x, v.y, v.z, 0}), qconj(q));
  return {r.x, r.y, r.z};
}
// q_a - q_b := 2 vec(q_a q_b^-1), canonicalized to w >= 0 (Math.swift:68-73)
template <typename T> V3<T> qsub(Q<T> a, Q<T> b) {
  Q<T> d = qmul(a, qinv(b));
  V3<T> v{d.x, d.y, d.z};
  return (d.w < 0 ? -v : v) * T(2);
}
// q + w := normalize(q + 1/2 (0, w) q) (Math.swift:75-79)
template <typename T> Q<T> qadd(Q<T> q, V3<T> w) {
  Q<T> dq = qmul(Q<T>{w.x, w.y, w.z, 0}, q);
  return qnorm(Q<T>{q.x + T(0.5) * dq.x, q.y + T(0.5) * dq.y, q.z + T(0.5) * dq.z, q.w + T(0.5) * dq.w});
}
template <typename T> V3<T> xform(V3<T> p, Q<T> q, V3<T> v) { return rot(q, v) + p; }

// row-major 3x3
template <typename T>
struct M3 {
  V3<T> r[3];
};
template <typename T> M3<T> mdiag(V3<T> d) {
  M3<T> m;
  m.r[0] = {d.x, 0, 0};
  m.r[1] = {0, d.y, 0};
  m.r[2] = {0, 0, d.z};
  return m;
}
template <typename T> M3<T> madd(const M3<T>& a, const M3<T>& b) {
  return {{a.r[0] + b.r[0], a.r[1] + b.r[1], a.r[2] + b.r[2]}};
}
template <typename T> M3<T> mscale(const M3<T>& a, T s) { return {{a.r[0] * s, a.r[1] * s, a.r[2] * s}}; }
template <typename T> M3<T> mtrans(const M3<T>& m) {
  M3<T> t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t.r[i][j] = m.r[j][i];
  return t;
}
template <typename T> M3<T> mmul(const M3<T>& a, const M3<T>& b) {
  M3<T> c;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c.r[i][j] = a.r[i][0] * b.r[0][j] + a.r[i][1] * b.r[1][j] + a.r[i][2] * b.r[2][j];
  return c;
}
template <typename T> V3<T> mvec(const M3<T>& m, V3<T> v) { return {dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)}; }
// row i = a[i] * b
template <typename T> M3<T> outer(V3<T> a, V3<T> b) { return {{b * a.x, b * a.y, b * a.z}}; }
// skew(r) v = r x v
template <typename T> M3<T> skew(V3<T> r) { return {{V3<T>(0, -r.z, r.y), V3<T>(r.z, 0, -r.x), V3<T>(-r.y, r.x, 0)}}; }
template <typename T> M3<T> world_inertia(Q<T> q, V3<T> m) {
  V3<T> c0 = rot(q, V3<T>(1, 0, 0)), c1 = rot(q, V3<T>(0, 1, 0)), c2 = rot(q, V3<T>(0, 0, 1));
  return madd(madd(mscale(outer(c0, c0), m.x), mscale(outer(c1, c1), m.y)), mscale(outer(c2, c2), m.z));
}
// diagonal of column norms (paper Sec 3.5)
template <typename T> M3<T> diagonalized(const M3<T>& m) {
  V3<T> c0{m.r[0].x, m.r[1].x, m.r[2].x}, c1{m.r[0].y, m.r[1].y, m.r[2].y}, c2{m.r[0].z, m.r[1].z, m.r[2].z};
  return mdiag(V3<T>(len(c0), len(c1), len(c2)));
}
template <typename T> M3<T> geom_ball_socket(int k, V3<T> v) {
  M3<T> m = mdiag(V3<T>(-v[k], -v[k], -v[k]));
  m.r[0][k] += v.x;
  m.r[1][k] += v.y;
  m.r[2][k] += v.z;
  return m;
}

// plain LDL^T of the symmetric 6x6 [aLin aCross^T; aCross aAng] (Math.swift:163-226)
template <typename T>
void solve6(const M3<T>& aLin, const M3<T>& aAng, const M3<T>& aCr, V3<T> bLin, V3<T> bAng, V3<T>& xLin,
            V3<T>& xAng) {
  T A[6][6];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      A[i][j] = aLin.r[i][j];
      A[3 + i][3 + j] = aAng.r[i][j];
      A[3 + i][j] = aCr.r[i][j];
      A[j][3 + i] = aCr.r[i][j];
    }
  T L[6][6] = {}, D[6];
  for (int j = 0; j < 6; j++) {
    T s = A[j][j];
    for (int k = 0; k < j; k++) s -= L[j][k] * L[j][k] * D[k];
    D[j] = s;
    for (int i = j + 1; i < 6; i++) {
      T t = A[i][j];
      for (int k = 0; k < j; k++) t -= L[j][k] * L[i][k] * D[k];
      L[i][j] = t / D[j];
    }
  }
  T b[6] = {bLin.x, bLin.y, bLin.z, bAng.x, bAng.y, bAng.z}, y[6], x[6];
  for (int i = 0; i < 6; i++) {
    T s = b[i];
    for (int k = 0; k < i; k++) s -= L[i][k] * y[k];
    y[i] = s;
  }
  for (int i = 5; i >= 0; i--) {
    T s = y[i] / D[i];
    for (int k = i + 1; k < 6; k++) s -= L[k][i] * x[k];
    x[i] = s;
  }
  xLin = {x[0], x[1], x[2]};
  xAng = {x[3], x[4], x[5]};
}

constexpr double kPenMin = 1.0, kPenMax = 1.0e10, kPenMaxT = 1.0e6, kStick = 1.0e-5;

template <typename T>
struct Body {
  V3<T> x, v, w, prevV, x0, xi, size, moment;
  Q<T> q, q0, qi;
  T mass = 0, friction = 0, dynFriction = 0, gravityScale = 1, radius = 0;
  uint32_t shape = 0;
  bool particle = false;
  std::vector<int> forces;  // indices into the solver's force table, in creation order
};

template <typename T>
struct Contact {
  int32_t key = 0;
  V3<T> rA, rB, C0, penalty, lambda;
  bool stick = false;
};

enum Kind { kJoint, kSpring, kManifold };

template <typename T>
struct Force {
  Kind kind;
  int a = -1, b = -1;  // -1 = world (joints only)
  bool alive = true;
  // joint
  V3<T> rA, rB, C0Lin, C0Ang, penLin, penAng, lamLin, lamAng;
  T stiffLin = 0, stiffAng = 0, fracture = 0, torqueArm = 0;
  bool broken = false, fractureLinear = false, hinge = false;
  V3<T> hingeAxis;
  Q<T> restRel;
  // spring
  T rest = 0, stiffness = 0, sLambda = 0, sPenalty = 0, sC0 = 0;
  bool hard = false;
  // manifold
  std::vector<Contact<T>> contacts;
  V3<T> n, t1, t2;
  T statF = 0, dynF = 0;
};

template <typename T>
struct CpuPath {
  std::vector<Body<T>> bodies;
  std::vector<Force<T>> forces;
  std::vector<int> order;  // live forces in creation order
  T dt, gravity, alpha, betaLin, betaAng, gamma, lambdaMax, margin, linDamp, angDamp;
  int iterations = 10;
  uint32_t frictionMode = 0;

  T combine(T a, T b) const {
    if (frictionMode == 1) return a * b;
    if (frictionMode == 2) return std::min(a, b);
    if (frictionMode == 3) return std::max(a, b);
    if (frictionMode == 4) return T(0.5) * (a + b);
    return std::sqrt(std::max(a * b, T(0)));
  }

  int init(const avbd_scene_desc* d) {
    const avbd_settings& s = d->settings;
    dt = s.dt;
    gravity = s.gravity;
    iterations = s.iterations;
    alpha = s.alpha;
    betaLin = s.beta_lin;
    betaAng = s.beta_ang;
    gamma = s.gamma;
    lambdaMax = s.lambda_max;
    margin = std::max((T)s.collision_margin, T(0));
    linDamp = s.rigid_linear_damping;
    angDamp = s.rigid_angular_damping;
    frictionMode = s.friction_combine_mode;
    if (d->num_tets || d->num_tris) return 4;  // the CPU path has rigid bodies only
    const uint32_t n = d->num_bodies;
    bodies.resize(n);
    for (uint32_t i = 0; i < n; i++) {
      Body<T>& b = bodies[i];
      b.size = {(T)d->body_size[3 * i], (T)d->body_size[3 * i + 1], (T)d->body_size[3 * i + 2]};
      b.shape = d->body_shape ? d->body_shape[i] : 0u;
      if (b.shape != 0u && b.shape != 1u) return 4;  // boxes and spheres
      const T rho = d->body_density[i];
      b.friction = d->body_friction[i];
      b.dynFriction = d->body_dynamic_friction ? d->body_dynamic_friction[i] : d->body_friction[i];
      b.gravityScale = d->body_gravity_scale ? d->body_gravity_scale[i] : 1.0f;
      b.particle = d->body_is_particle && d->body_is_particle[i];
      b.x = {(T)d->body_position[3 * i], (T)d->body_position[3 * i + 1], (T)d->body_position[3 * i + 2]};
      if (d->body_rotation)
        b.q = {(T)d->body_rotation[4 * i], (T)d->body_rotation[4 * i + 1], (T)d->body_rotation[4 * i + 2],
               (T)d->body_rotation[4 * i + 3]};
      if (d->body_velocity)
        b.v = {(T)d->body_velocity[3 * i], (T)d->body_velocity[3 * i + 1], (T)d->body_velocity[3 * i + 2]};
      b.prevV = b.v;
      // mass / moment per shape (CPUSolver.swift:56-95)
      if (b.shape == 0u) {
        T m = b.size.x * b.size.y * b.size.z * rho;
        b.mass = rho > 0 ? m : T(0);
        b.moment = {(b.size.y * b.size.y + b.size.z * b.size.z) / 12 * m,
                    (b.size.x * b.size.x + b.size.z * b.size.z) / 12 * m,
                    (b.size.x * b.size.x + b.size.y * b.size.y) / 12 * m};
        b.radius = len(b.size * T(0.5));
      } else {
        T r = b.size.x / 2;
        T m = rho > 0 ? T(4.0 / 3.0) * T(M_PI) * r * r * r * rho : T(0);
        b.mass = m;
        b.moment = {T(0.4) * m * r * r, T(0.4) * m * r * r, T(0.4) * m * r * r};
        b.radius = r;
      }
      if (d->body_mass && !std::isnan(d->body_mass[i])) {
        b.mass = d->body_mass[i];
        b.moment = {(T)d->body_inertia[3 * i], (T)d->body_inertia[3 * i + 1], (T)d->body_inertia[3 * i + 2]};
      }
    }
    for (uint32_t j = 0; j < d->num_joints; j++) {
      Force<T> f;
      f.kind = kJoint;
      f.a = d->joint_body_a[j];
      f.b = d->joint_body_b[j];
      f.rA = {(T)d->joint_ra[3 * j], (T)d->joint_ra[3 * j + 1], (T)d->joint_ra[3 * j + 2]};
      f.rB = {(T)d->joint_rb[3 * j], (T)d->joint_rb[3 * j + 1], (T)d->joint_rb[3 * j + 2]};
      f.stiffLin = d->joint_stiffness_lin[j];
      f.stiffAng = d->joint_stiffness_ang[j];
      f.fracture = d->joint_fracture ? (T)d->joint_fracture[j] : std::numeric_limits<T>::infinity();
      f.fractureLinear = d->joint_fracture_linear && d->joint_fracture_linear[j];
      V3<T> sa = f.a >= 0 ? bodies[f.a].size : V3<T>();
      f.torqueArm = dot(sa + bodies[f.b].size, sa + bodies[f.b].size);
      if (f.stiffLin > 0 && std::isfinite(f.stiffLin)) f.penLin = V3<T>(1, 1, 1) * std::min(f.stiffLin, T(1e9));
      if (f.stiffAng > 0 && std::isfinite(f.stiffAng)) f.penAng = V3<T>(1, 1, 1) * std::min(f.stiffAng, T(1e9));
      Q<T> qA0 = f.a >= 0 ? bodies[f.a].q : Q<T>();
      f.restRel = qnorm(qmul(qinv(qA0), bodies[f.b].q));
      if (d->joint_hinge_axis && !std::isnan(d->joint_hinge_axis[3 * j])) {
        f.hinge = true;
        f.hingeAxis = {(T)d->joint_hinge_axis[3 * j], (T)d->joint_hinge_axis[3 * j + 1],
                       (T)d->joint_hinge_axis[3 * j + 2]};
      }
      add_force(f);
    }
    for (uint32_t k = 0; k < d->num_springs; k++) {
      Force<T> f;
      f.kind = kSpring;
      f.a = (int)d->spring_body_a[k];
      f.b = (int)d->spring_body_b[k];
      f.rA = {(T)d->spring_ra[3 * k], (T)d->spring_ra[3 * k + 1], (T)d->spring_ra[3 * k + 2]};
      f.rB = {(T)d->spring_rb[3 * k], (T)d->spring_rb[3 * k + 1], (T)d->spring_rb[3 * k + 2]};
      f.stiffness = d->spring_stiffness[k];
      T rest = d->spring_rest[k];
      if (rest < 0) {
        V3<T> pA = xform(bodies[f.a].x, bodies[f.a].q, f.rA), pB = xform(bodies[f.b].x, bodies[f.b].q, f.rB);
        rest = len(pA - pB);
      }
      f.rest = rest;
      f.hard = d->spring_hard && d->spring_hard[k];
      add_force(f);
    }
    return 0;
  }

  void add_force(const Force<T>& f) {
    const int id = (int)forces.size();
    forces.push_back(f);
    order.push_back(id);
    if (f.a >= 0) bodies[f.a].forces.push_back(id);
    if (f.b >= 0) bodies[f.b].forces.push_back(id);
  }
  void drop_force(int id) {
    Force<T>& f = forces[id];
    f.alive = false;
    for (int bi : {f.a, f.b})
      if (bi >= 0) {
        auto& v = bodies[bi].forces;
        v.erase(std::remove(v.begin(), v.end(), id), v.end());
      }
  }
  bool constrained(int i, int j) const {
    for (int id : bodies[i].forces) {
      const Force<T>& f = forces[id];
      if ((f.a == i && f.b == j) || (f.a == j && f.b == i)) return true;
    }
    return false;
  }

  // ---- collision (CPUCollide.swift) ----
  struct Obb {
    V3<T> c, h, ax[3];
  };
  Obb obb(const Body<T>& b) const {
    Obb o;
    o.c = b.x;
    o.h = b.size * T(0.5);
    o.ax[0] = rot(b.q, V3<T>(1, 0, 0));
    o.ax[1] = rot(b.q, V3<T>(0, 1, 0));
    o.ax[2] = rot(b.q, V3<T>(0, 0, 1));
    return o;
  }
  struct Sat {
    int type = 0, ia = -1, ib = -1;
    T sep = -std::numeric_limits<T>::max();
    V3<T> n;
    bool valid = false;
  };
  static bool test_axis(const Obb& A, const Obb& B, V3<T> delta, V3<T> axis, int type, int ia, int ib, Sat& best) {
    T l2 = dot(axis, axis);
    if (l2 < T(1e-6)) return true;
    V3<T> n = axis / std::sqrt(l2);
    if (dot(n, delta) < 0) n = -n;
    T dist = std::fabs(dot(delta, n));
    T rA = A.h.x * std::fabs(dot(n, A.ax[0])) + A.h.y * std::fabs(dot(n, A.ax[1])) + A.h.z * std::fabs(dot(n, A.ax[2]));
    T rB = B.h.x * std::fabs(dot(n, B.ax[0])) + B.h.y * std::fabs(dot(n, B.ax[1])) + B.h.z * std::fabs(dot(n, B.ax[2]));
    T sep = dist - (rA + rB);
    if (sep > 0) return false;
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
  static V3<T> support(const Obb& b, V3<T> d) {
    V3<T> p = b.c;
    for (int k = 0; k < 3; k++) p = p + b.ax[k] * (b.h[k] * (dot(d, b.ax[k]) >= 0 ? T(1) : T(-1)));
    return p;
  }
  static std::vector<V3<T>> clip(const std::vector<V3<T>>& in, V3<T> pn, T off) {
    std::vector<V3<T>> out;
    if (in.empty()) return out;
    V3<T> a = in.back();
    T da = dot(pn, a) - off;
    for (const V3<T>& b : in) {
      T db = dot(pn, b) - off;
      bool ai = da <= T(1e-5), bi = db <= T(1e-5);
      if (ai != bi) {
        T t = 0, den = da - db;
        if (std::fabs(den) > T(1e-6)) t = std::min(std::max(da / den, T(0)), T(1));
        if (out.size() < 16) out.push_back(a + (b - a) * t);
      }
      if (bi && out.size() < 16) out.push_back(b);
      a = b;
      da = db;
    }
    return out;
  }
  static void seg_seg(V3<T> p0, V3<T> p1, V3<T> q0, V3<T> q1, V3<T>& c0, V3<T>& c1) {
    V3<T> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), s = 0, t = 0;
    auto cl = [](T v) { return std::min(std::max(v, T(0)), T(1)); };
    if (a <= T(1e-6) && e <= T(1e-6)) {
      c0 = p0;
      c1 = q0;
      return;
    }
    if (a <= T(1e-6)) {
      t = cl(f / e);
    } else {
      T c = dot(d1, r);
      if (e <= T(1e-6)) {
        s = cl(-c / a);
      } else {
        T bb = dot(d1, d2), den = a * e - bb * bb;
        if (std::fabs(den) > T(1e-6)) s = cl((bb * f - c * e) / den);
        t = (bb * s + f) / e;
        if (t < 0) {
          t = 0;
          s = cl(-c / a);
        } else if (t > 1) {
          t = 1;
          s = cl((bb - c) / a);
        }
      }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
  }
  static void face_axes(const Obb& b, int k, V3<T>& u, V3<T>& v, T& eu, T& ev) {
    int i = k == 0 ? 1 : 0, j = k == 2 ? 1 : 2;
    u = b.ax[i];
    v = b.ax[j];
    eu = b.h[i];
    ev = b.h[j];
  }
  // contact anchors: body-local for boxes, world offsets for spheres
  V3<T> anchor(const Body<T>& b, V3<T> xw) const {
    return b.shape != 0u ? xw - b.x : rot(qconj(b.q), xw - b.x);
  }
  V3<T> anchor_world(const Body<T>& b, V3<T> r) const { return b.shape != 0u ? b.x + r : xform(b.x, b.q, r); }
  V3<T> anchor_offset(const Body<T>& b, V3<T> r) const { return b.shape != 0u ? r : rot(b.q, r); }

  void add_contact(const Body<T>& A, const Body<T>& B, std::vector<Contact<T>>& cs, std::vector<V3<T>>& mids,
                   V3<T> xA, V3<T> xB, int32_t key) const {
    V3<T> mid = (xA + xB) * T(0.5);
    for (const V3<T>& m : mids)
      if (dot(mid - m, mid - m) < T(1e-6)) return;
    if (cs.size() >= 8) return;
    Contact<T> c;
    c.key = key;
    c.rA = rot(qconj(A.q), xA - A.x);
    c.rB = rot(qconj(B.q), xB - B.x);
    cs.push_back(c);
    mids.push_back(mid);
  }
  static void basis_of(V3<T> n, V3<T>& t1, V3<T>& t2) {
    t1 = std::fabs(n.x) > std::fabs(n.z) ? V3<T>(-n.y, n.x, 0) : V3<T>(0, -n.z, n.y);
    t1 = t1 / len(t1);
    t2 = cross(n, t1);
  }
  // returns the contacts and the normal (B toward A)
  bool collide(int ia, int ib, std::vector<Contact<T>>& cs, V3<T>& nrm) const {
    const Body<T>& A = bodies[ia];
    const Body<T>& B = bodies[ib];
    cs.clear();
    if (A.shape == 1u || B.shape == 1u) {
      Contact<T> c;
      if (A.shape == 1u && B.shape == 1u) {
        T rA = A.size.x / 2, rB = B.size.x / 2;
        V3<T> d = A.x - B.x;
        T dist = len(d);
        if (dist > rA + rB + margin) return false;
        nrm = dist > T(1e-9) ? d / dist : V3<T>(0, 0, 1);
        c.rA = (A.x - nrm * rA) - A.x;
        c.rB = (B.x + nrm * rB) - B.x;
        cs.push_back(c);
        return true;
      }
      const bool sA = A.shape == 1u;
      const Body<T>& S = sA ? A : B;
      const Body<T>& X = sA ? B : A;
      T r = S.size.x / 2;
      V3<T> h = X.size * T(0.5);
      V3<T> loc = rot(qconj(X.q), S.x - X.x);
      V3<T> cl{std::min(std::max(loc.x, -h.x), h.x), std::min(std::max(loc.y, -h.y), h.y),
               std::min(std::max(loc.z, -h.z), h.z)};
      V3<T> nL, qL;
      if (cl.x == loc.x && cl.y == loc.y && cl.z == loc.z) {
        V3<T> pen = h - vabs(loc);
        int ax = 0;
        if (pen.y < pen[ax]) ax = 1;
        if (pen.z < pen[ax]) ax = 2;
        nL = V3<T>();
        nL[ax] = loc[ax] >= 0 ? T(1) : T(-1);
        qL = loc;
        qL[ax] = nL[ax] * h[ax];
      } else {
        V3<T> d = loc - cl;
        T dist = len(d);
        if (dist > r + margin) return false;
        nL = d / std::max(dist, T(1e-9));
        qL = cl;
      }
      V3<T> nW = rot(X.q, nL), qW = xform(X.x, X.q, qL), xS = S.x - nW * r;
      nrm = sA ? nW : -nW;
      V3<T> xA = sA ? xS : qW, xB = sA ? qW : xS;
      c.rA = anchor(A, xA);
      c.rB = anchor(B, xB);
      cs.push_back(c);
      return true;
    }
    // box vs box: SAT, reference-face clipping or the edge pair
    Obb a = obb(A), b = obb(B);
    V3<T> delta = b.c - a.c;
    Sat face, edge;
    for (int i = 0; i < 3; i++)
      if (!test_axis(a, b, delta, a.ax[i], 0, i, -1, face)) return false;
    for (int i = 0; i < 3; i++)
      if (!test_axis(a, b, delta, b.ax[i], 1, -1, i, face)) return false;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        if (!test_axis(a, b, delta, cross(a.ax[i], b.ax[j]), 2, i, j, edge)) return false;
    if (!face.valid) return false;
    Sat best = face;
    if (edge.valid && T(0.95) * edge.sep > face.sep + T(0.01)) best = edge;
    nrm = -best.n;
    std::vector<V3<T>> mids;
    if (best.type == 2) {
      auto edge_of = [](const Obb& o, int k, V3<T> dir, V3<T>& e0, V3<T>& e1) {
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        V3<T> c = o.c + o.ax[k1] * (o.h[k1] * (dot(dir, o.ax[k1]) >= 0 ? T(1) : T(-1))) +
                  o.ax[k2] * (o.h[k2] * (dot(dir, o.ax[k2]) >= 0 ? T(1) : T(-1)));
        e0 = c - o.ax[k] * o.h[k];
        e1 = c + o.ax[k] * o.h[k];
      };
      V3<T> a0, a1, b0, b1, xA, xB;
      edge_of(a, best.ia, best.n, a0, a1);
      edge_of(b, best.ib, -best.n, b0, b1);
      seg_seg(a0, a1, b0, b1, xA, xB);
      int32_t key = (2 << 24) | ((best.ia & 0xFF) << 8) | (best.ib & 0xFF);
      add_contact(A, B, cs, mids, xA, xB, key);
      if (cs.empty()) add_contact(A, B, cs, mids, support(a, best.n), support(b, -best.n), key);
      return true;
    }
    const bool refA = best.type == 0;
    const int ra = refA ? best.ia : best.ib;
    const Obb& R = refA ? a : b;
    const Obb& I = refA ? b : a;
    V3<T> out = refA ? best.n : -best.n;
    V3<T> rN = R.ax[ra] * (dot(out, R.ax[ra]) >= 0 ? T(1) : T(-1));
    V3<T> rC = R.c + rN * R.h[ra];
    V3<T> rU, rV, iU, iV;
    T eU, eV, fU, fV;
    face_axes(R, ra, rU, rV, eU, eV);
    int inc = 0;
    T bd = -std::numeric_limits<T>::max();
    for (int i = 0; i < 3; i++) {
      T dd = std::fabs(dot(I.ax[i], rN));
      if (dd > bd) {
        bd = dd;
        inc = i;
      }
    }
    V3<T> iN = I.ax[inc] * (dot(I.ax[inc], rN) > 0 ? T(-1) : T(1));
    V3<T> iC = I.c + iN * I.h[inc];
    face_axes(I, inc, iU, iV, fU, fV);
    std::vector<V3<T>> poly = {iC + iU * fU + iV * fV, iC - iU * fU + iV * fV, iC - iU * fU - iV * fV,
                               iC + iU * fU - iV * fV};
    poly = clip(poly, rU, dot(rU, rC) + eU);
    if (!poly.empty()) poly = clip(poly, -rU, dot(-rU, rC) + eU);
    if (!poly.empty()) poly = clip(poly, rV, dot(rV, rC) + eV);
    if (!poly.empty()) poly = clip(poly, -rV, dot(-rV, rC) + eV);
    int32_t prefix = ((refA ? 0 : 1) << 24) | ((ra & 0xFF) << 16) | ((inc & 0xFF) << 8);
    for (size_t i = 0; i < poly.size() && cs.size() < 8; i++) {
      T dd = dot(poly[i] - rC, rN);
      if (dd > T(1e-5)) continue;
      V3<T> pR = poly[i] - rN * dd;
      add_contact(A, B, cs, mids, refA ? pR : poly[i], refA ? poly[i] : pR, prefix | (int32_t)(i & 0xFF));
    }
    if (cs.empty() && !poly.empty())
      add_contact(A, B, cs, mids, support(a, best.n), support(b, -best.n), prefix);
    return !cs.empty();
  }

  // ---- forces ----
  bool init_manifold(Force<T>& f) {
    const Body<T>& A = bodies[f.a];
    const Body<T>& B = bodies[f.b];
    f.statF = combine(A.friction, B.friction);
    f.dynF = combine(A.dynFriction, B.dynFriction);
    std::vector<Contact<T>> fresh;
    V3<T> n;
    if (!collide(f.a, f.b, fresh, n)) fresh.clear();
    else {
      f.n = n;
      basis_of(n, f.t1, f.t2);
    }
    const bool round = A.shape != 0u || B.shape != 0u;
    for (Contact<T>& c : fresh)
      for (const Contact<T>& old : f.contacts)
        if (old.key == c.key) {
          V3<T> ra = c.rA, rb = c.rB;
          c = old;
          if (!old.stick || round) {
            c.rA = ra;
            c.rB = rb;
          }
          break;
        }
    f.contacts = fresh;
    T mm = std::numeric_limits<T>::max();
    if (A.mass > 0) mm = std::min(mm, A.mass);
    if (B.mass > 0) mm = std::min(mm, B.mass);
    if (mm == std::numeric_limits<T>::max()) mm = 0;
    T kf = std::min(T(kPenMax), std::max(T(kPenMin), mm / std::max(dt * dt, T(1e-12))));
    V3<T> pmin{kf, std::min(kf, T(kPenMaxT)), std::min(kf, T(kPenMaxT))};
    for (Contact<T>& c : f.contacts) {
      V3<T> d = anchor_world(A, c.rA) - anchor_world(B, c.rB);
      c.C0 = {dot(f.n, d) + margin, dot(f.t1, d), dot(f.t2, d)};
      c.lambda = c.lambda * (alpha * gamma);
      V3<T> p = c.penalty * gamma;
      c.penalty = {std::min(std::max(p.x, pmin.x), T(kPenMax)), std::min(std::max(p.y, pmin.y), T(kPenMaxT)),
                   std::min(std::max(p.z, pmin.z), T(kPenMaxT))};
    }
    return !f.contacts.empty();
  }
  V3<T> joint_clin(const Force<T>& f) const {
    V3<T> pA = f.a >= 0 ? xform(bodies[f.a].x, bodies[f.a].q, f.rA) : f.rA;
    return pA - xform(bodies[f.b].x, bodies[f.b].q, f.rB);
  }
  V3<T> joint_cang(const Force<T>& f) const {
    Q<T> qA = f.a >= 0 ? bodies[f.a].q : Q<T>();
    Q<T> qAr = qnorm(qmul(qA, f.restRel));
    if (f.hinge) return cross(rot(qAr, f.hingeAxis), rot(bodies[f.b].q, f.hingeAxis)) * f.torqueArm;
    return qsub(qAr, bodies[f.b].q) * f.torqueArm;
  }
  bool init_force(Force<T>& f) {
    if (f.kind == kManifold) return init_manifold(f);
    if (f.kind == kJoint) {
      f.C0Lin = joint_clin(f);
      f.C0Ang = joint_cang(f);
      f.lamLin = f.lamLin * (alpha * gamma);
      f.lamAng = f.lamAng * (alpha * gamma);
      auto warm = [&](V3<T> p, T stiff) {
        V3<T> q = p * gamma;
        for (int k = 0; k < 3; k++) q[k] = std::min(std::min(std::max(q[k], T(kPenMin)), T(kPenMax)), stiff);
        return q;
      };
      f.penLin = warm(f.penLin, f.stiffLin);
      f.penAng = warm(f.penAng, f.stiffAng);
      return !f.broken;
    }
    if (f.hard) {
      V3<T> d = xform(bodies[f.a].x, bodies[f.a].q, f.rA) - xform(bodies[f.b].x, bodies[f.b].q, f.rB);
      f.sC0 = len(d) - f.rest;
      f.sLambda *= alpha * gamma;
      f.sPenalty = std::min(std::min(std::max(f.sPenalty * gamma, T(kPenMin)), T(kPenMax)), f.stiffness);
    }
    return true;
  }

  // Taylor-linearized contact C and the clamped force (CPUManifold.swift:123-150)
  void contact_fc(const Force<T>& f, const Contact<T>& c, V3<T> dAL, V3<T> dAA, V3<T> dBL, V3<T> dBA, V3<T> rAW,
                  V3<T> rBW, V3<T>& F, V3<T>& C, T& fs, T& bound) const {
    const V3<T> j[3] = {f.n, f.t1, f.t2};
    C = c.C0 * (T(1) - alpha);
    for (int k = 0; k < 3; k++)
      C[k] = C[k] + dot(j[k], dAL) - dot(j[k], dBL) + dot(cross(rAW, j[k]), dAA) + dot(cross(rBW, -j[k]), dBA);
    F = mulc(c.penalty, C) + c.lambda;
    F.x = std::min(F.x, T(0));
    bound = std::fabs(F.x) * f.statF;
    const T dynB = std::fabs(F.x) * f.dynF;
    fs = std::sqrt(F.y * F.y + F.z * F.z);
    if (fs > bound && fs > 0) {
      F.y = F.y * (dynB / fs);
      F.z = F.z * (dynB / fs);
    }
  }

  void primal_force(const Force<T>& f, int self, M3<T>& hL, M3<T>& hA, M3<T>& hC, V3<T>& gL, V3<T>& gA) const {
    const bool isA = self == f.a;
    if (f.kind == kManifold) {
      const Body<T>& A = bodies[f.a];
      const Body<T>& B = bodies[f.b];
      V3<T> dAL = A.x - A.x0, dAA = qsub(A.q, A.q0), dBL = B.x - B.x0, dBA = qsub(B.q, B.q0);
      for (const Contact<T>& c : f.contacts) {
        V3<T> rAW = anchor_offset(A, c.rA), rBW = anchor_offset(B, c.rB), F, C;
        T fs, bd;
        contact_fc(f, c, dAL, dAA, dBL, dBA, rAW, rBW, F, C, fs, bd);
        T s = isA ? T(1) : T(-1);
        M3<T> jL{{f.n * s, f.t1 * s, f.t2 * s}};
        V3<T> rW = isA ? rAW : rBW;
        M3<T> jA{{cross(rW, jL.r[0]), cross(rW, jL.r[1]), cross(rW, jL.r[2])}};
        M3<T> K = mdiag(c.penalty);
        M3<T> jLT = mtrans(jL), jAT = mtrans(jA), jATk = mmul(jAT, K);
        hL = madd(hL, mmul(mmul(jLT, K), jL));
        hA = madd(hA, mmul(jATk, jA));
        hC = madd(hC, mmul(jATk, jL));
        gL = gL + mvec(jLT, F);
        gA = gA + mvec(jAT, F);
      }
      return;
    }
    if (f.kind == kJoint) {
      const Body<T>& B = bodies[f.b];
      if (dot(f.penLin, f.penLin) > 0) {
        V3<T> C = joint_clin(f);
        if (std::isinf(f.stiffLin)) C = C - f.C0Lin * alpha;
        V3<T> F = mulc(f.penLin, C) + f.lamLin;
        M3<T> jL = isA ? mdiag(V3<T>(1, 1, 1)) : mdiag(V3<T>(-1, -1, -1));
        M3<T> jA = isA ? skew(-rot(bodies[f.a].q, f.rA)) : skew(rot(B.q, f.rB));
        M3<T> jLT = mtrans(jL), jAT = mtrans(jA), K = mdiag(f.penLin), jATk = mmul(jAT, K);
        hL = madd(hL, mmul(mmul(jLT, K), jL));
        hA = madd(hA, mmul(jATk, jA));
        hC = madd(hC, mmul(jATk, jL));
        V3<T> r = isA ? rot(bodies[f.a].q, f.rA) : -rot(B.q, f.rB);
        M3<T> H = madd(madd(mscale(geom_ball_socket(0, r), F.x), mscale(geom_ball_socket(1, r), F.y)),
                       mscale(geom_ball_socket(2, r), F.z));
        hA = madd(hA, diagonalized(H));
        gL = gL + mvec(jLT, F);
        gA = gA + mvec(jAT, F);
      }
      if (dot(f.penAng, f.penAng) > 0) {
        V3<T> C = joint_cang(f);
        if (std::isinf(f.stiffAng)) C = C - f.C0Ang * alpha;
        V3<T> F = mulc(f.penAng, C) + f.lamAng;
        T s = (isA ? T(1) : T(-1)) * f.torqueArm;
        if (f.hinge) s = -s;
        hA = madd(hA, mdiag(f.penAng * (s * s)));
        gA = gA + F * s;
      }
      return;
    }
    const Body<T>& A = bodies[f.a];
    const Body<T>& B = bodies[f.b];
    V3<T> d = xform(A.x, A.q, f.rA) - xform(B.x, B.q, f.rB);
    T dl = len(d);
    if (dl <= T(1e-6)) return;
    V3<T> n = d / dl;
    T k = f.stiffness, fm;
    if (f.hard) {
      k = f.sPenalty;
      fm = f.sPenalty * (dl - f.rest - f.sC0 * alpha) + f.sLambda;
    } else {
      fm = f.stiffness * (dl - f.rest);
    }
    V3<T> rW = isA ? rot(A.q, f.rA) : rot(B.q, f.rB);
    V3<T> jL = isA ? n : -n;
    V3<T> jA = isA ? cross(rW, n) : -cross(rW, n);
    hL = madd(hL, mscale(outer(jL, jL), k));
    hA = madd(hA, mscale(outer(jA, jA), k));
    hC = madd(hC, mscale(outer(jA, jL), k));
    gL = gL + jL * fm;
    gA = gA + jA * fm;
  }

  void dual_force(Force<T>& f) {
    if (f.kind == kManifold) {
      const Body<T>& A = bodies[f.a];
      const Body<T>& B = bodies[f.b];
      V3<T> dAL = A.x - A.x0, dAA = qsub(A.q, A.q0), dBL = B.x - B.x0, dBA = qsub(B.q, B.q0);
      for (Contact<T>& c : f.contacts) {
        V3<T> rAW = anchor_offset(A, c.rA), rBW = anchor_offset(B, c.rB), F, C;
        T fs, bd;
        contact_fc(f, c, dAL, dAA, dBL, dBA, rAW, rBW, F, C, fs, bd);
        V3<T> Fb = F;
        Fb.x = std::max(Fb.x, -lambdaMax);
        c.lambda = Fb;
        if (F.x < 0) c.penalty.x = std::min(c.penalty.x + betaLin * std::fabs(C.x), T(kPenMax));
        if (fs <= bd) {
          c.penalty.y = std::min(c.penalty.y + betaLin * std::fabs(C.y), T(kPenMaxT));
          c.penalty.z = std::min(c.penalty.z + betaLin * std::fabs(C.z), T(kPenMaxT));
          c.stick = std::sqrt(C.y * C.y + C.z * C.z) < T(kStick);
        }
      }
      return;
    }
    if (f.kind == kJoint) {
      if (dot(f.penLin, f.penLin) > 0) {
        V3<T> C = joint_clin(f);
        if (std::isinf(f.stiffLin)) {
          C = C - f.C0Lin * alpha;
          f.lamLin = mulc(f.penLin, C) + f.lamLin;
        }
        T cap = std::min(f.stiffLin, T(kPenMax));
        for (int k = 0; k < 3; k++) f.penLin[k] = std::min(f.penLin[k] + std::fabs(C[k]) * betaLin, cap);
      }
      if (dot(f.penAng, f.penAng) > 0) {
        V3<T> C = joint_cang(f);
        if (std::isinf(f.stiffAng)) {
          C = C - f.C0Ang * alpha;
          f.lamAng = mulc(f.penAng, C) + f.lamAng;
        }
        T cap = std::min(f.stiffAng, T(kPenMax));
        for (int k = 0; k < 3; k++) f.penAng[k] = std::min(f.penAng[k] + std::fabs(C[k]) * betaAng, cap);
      }
      T lin2 = f.fractureLinear ? dot(f.lamLin, f.lamLin) : T(0);
      if (dot(f.lamAng, f.lamAng) + lin2 > f.fracture * f.fracture) {
        f.penLin = f.penAng = f.lamLin = f.lamAng = V3<T>();
        f.broken = true;
      }
      return;
    }
    if (!f.hard) return;
    V3<T> d = xform(bodies[f.a].x, bodies[f.a].q, f.rA) - xform(bodies[f.b].x, bodies[f.b].q, f.rB);
    T C = len(d) - f.rest - f.sC0 * alpha;
    f.sLambda = std::min(std::max(f.sPenalty * C + f.sLambda, -lambdaMax), lambdaMax);
    f.sPenalty = std::min(f.sPenalty + std::fabs(C) * betaLin, std::min(f.stiffness, T(kPenMax)));
  }

  // largest joint position error / contact penetration (CPUSolver.swift:318-331)
  T max_error() const {
    T err = 0;
    for (int id : order) {
      const Force<T>& f = forces[id];
      if (f.kind == kJoint) {
        if (f.stiffLin == 0 && f.stiffAng == 0) continue;
        err = std::max(err, len(joint_clin(f)));
      } else if (f.kind == kManifold) {
        for (const Contact<T>& c : f.contacts) {
          T pen = dot(f.n, anchor_world(bodies[f.a], c.rA) - anchor_world(bodies[f.b], c.rB)) + margin;
          err = std::max(err, std::max(T(0), -pen));
        }
      }
    }
    return err;
  }

  // CPUSolver.step (CPUSolver.swift:211-315)
  void step() {
    const int n = (int)bodies.size();
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) {
        V3<T> dp = bodies[i].x - bodies[j].x;
        T r = bodies[i].radius + bodies[j].radius;
        if (dot(dp, dp) <= r * r && !constrained(i, j)) {
          Force<T> m;
          m.kind = kManifold;
          m.a = i;
          m.b = j;
          add_force(m);
        }
      }
    std::vector<int> keep;
    for (int id : order) {
      if (init_force(forces[id])) keep.push_back(id);
      else drop_force(id);
    }
    order = keep;
    for (Body<T>& b : bodies) {
      b.xi = b.x + b.v * dt;
      if (b.mass > 0) b.xi = b.xi + V3<T>(0, 0, gravity * b.gravityScale) * (dt * dt);
      b.qi = qadd(b.q, b.w * dt);
      V3<T> acc = (b.v - b.prevV) / dt;
      T ext = acc.z * (gravity < 0 ? T(-1) : (gravity > 0 ? T(1) : T(0)));
      T wgt = std::min(std::max(ext / std::fabs(gravity), T(0)), T(1));
      if (!std::isfinite(wgt)) wgt = 0;
      b.x0 = b.x;
      b.q0 = b.q;
      if (b.mass > 0) {
        b.x = b.x + b.v * dt + V3<T>(0, 0, gravity * b.gravityScale) * (wgt * dt * dt);
        b.q = qadd(b.q, b.w * dt);
      }
    }
    for (int it = 0; it < iterations; it++) {
      for (int i = 0; i < n; i++) {
        Body<T>& b = bodies[i];
        if (!(b.mass > 0)) continue;
        const T idt2 = T(1) / (dt * dt);
        M3<T> hL = mdiag(V3<T>(1, 1, 1) * (b.mass * idt2));
        M3<T> inertia = mscale(world_inertia(b.q, b.moment), idt2);
        M3<T> hA = inertia, hC{};
        V3<T> gL = (b.x - b.xi) * (b.mass * idt2);
        V3<T> gA = mvec(inertia, qsub(b.q, b.qi));
        for (int id : b.forces) primal_force(forces[id], i, hL, hA, hC, gL, gA);
        if (b.particle) {
          hA = mdiag(V3<T>(1, 1, 1));
          hC = M3<T>{};
          gA = V3<T>();
        }
        V3<T> dL, dA;
        solve6(hL, hA, hC, -gL, -gA, dL, dA);
        b.x = b.x + dL;
        b.q = qadd(b.q, dA);
      }
      for (int id : order) dual_force(forces[id]);
    }
    for (Body<T>& b : bodies) {
      b.prevV = b.v;
      if (b.mass > 0) {
        b.v = (b.x - b.x0) / dt;
        b.w = qsub(b.q, b.q0) / dt;
        if (!b.particle) {
          b.v = b.v * std::exp(-linDamp * dt);
          b.w = b.w * std::exp(-angDamp * dt);
        }
      }
    }
  }
};

}  // namespace

struct oracle_cpu_solver {
  CpuPath<float>* f = nullptr;
  CpuPath<double>* d = nullptr;
};

extern "C" {

int oracle_cpu_create(const avbd_scene_desc* scene, int precision, oracle_cpu_solver** out) {
  if (!scene || !out) return 1;
  oracle_cpu_solver* o = new oracle_cpu_solver();
  int rc;
  if (precision == 64) rc = (o->d = new CpuPath<double>())->init(scene);
  else rc = (o->f = new CpuPath<float>())->init(scene);
  if (rc) {
    oracle_cpu_destroy(o);
    *out = nullptr;
    return rc;
  }
  *out = o;
  return 0;
}

void oracle_cpu_destroy(oracle_cpu_solver* o) {
  if (!o) return;
  delete o->f;
  delete o->d;
  delete o;
}

int oracle_cpu_step(oracle_cpu_solver* o, int n) {
  for (int k = 0; k < n; k++) {
    if (o->f) o->f->step();
    else o->d->step();
  }
  return 0;
}

int oracle_cpu_get_body_states_f64(oracle_cpu_solver* o, uint32_t n, double* pos3, double* rot4, double* vel3,
                                   double* angvel3) {
  auto go = [&](auto* s) {
    for (uint32_t i = 0; i < n && i < s->bodies.size(); i++) {
      const auto& b = s->bodies[i];
      if (pos3) { pos3[3 * i] = b.x.x; pos3[3 * i + 1] = b.x.y; pos3[3 * i + 2] = b.x.z; }
      if (rot4) { rot4[4 * i] = b.q.x; rot4[4 * i + 1] = b.q.y; rot4[4 * i + 2] = b.q.z; rot4[4 * i + 3] = b.q.w; }
      if (vel3) { vel3[3 * i] = b.v.x; vel3[3 * i + 1] = b.v.y; vel3[3 * i + 2] = b.v.z; }
      if (angvel3) { angvel3[3 * i] = b.w.x; angvel3[3 * i + 1] = b.w.y; angvel3[3 * i + 2] = b.w.z; }
    }
  };
  if (o->f) go(o->f);
  else go(o->d);
  return 0;
}

double oracle_cpu_max_constraint_error(oracle_cpu_solver* o) {
  return o->f ? (double)o->f->max_error() : o->d->max_error();
}

int oracle_cpu_num_manifolds(oracle_cpu_solver* o) {
  auto go = [&](auto* s) {
    int m = 0;
    for (int id : s->order) m += s->forces[id].kind == kManifold;
    return m;
  };
  return o->f ? go(o->f) : go(o->d);
}

}  // extern "C"
