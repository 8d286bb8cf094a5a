// stamps.hpp -- per-element gradient / Hessian-block contributions to one
// body's local Newton system (40_solver.metal:724-1360).
//
// A body's system is  [Hll Hla^T; Hla Haa] [dx; dth] = -[gl; ga].
// Acc keeps the lower triangle of Hll and Haa (all the LDL^T reads), the full
// angular-by-linear block Hla, and the two right-hand sides.
#pragma once
#include "state.hpp"

namespace avbd {

struct Acc {
  float l00, l10, l11, l20, l21, l22;  // Hll lower triangle
  float a00, a10, a11, a20, a21, a22;  // Haa lower triangle
  M3 cr;                               // Hla (rows angular, columns linear)
  V3 gl, ga;                           // gradients
};

AVBD_HD void acc_zero(Acc& a) {
  a.l00 = a.l10 = a.l11 = a.l20 = a.l21 = a.l22 = 0.0f;
  a.a00 = a.a10 = a.a11 = a.a20 = a.a21 = a.a22 = 0.0f;
  a.cr = m3_zero();
  a.gl = v3s(0);
  a.ga = v3s(0);
}
// Hll += s * u u^T
AVBD_HD void acc_lin_outer(Acc& a, V3 u, float s) {
  V3 su = u * s;
  a.l00 += su.x * u.x; a.l10 += su.y * u.x; a.l11 += su.y * u.y;
  a.l20 += su.z * u.x; a.l21 += su.z * u.y; a.l22 += su.z * u.z;
}
AVBD_HD void acc_lin_diag(Acc& a, float d) { a.l00 += d; a.l11 += d; a.l22 += d; }
AVBD_HD void acc_lin_m3(Acc& a, const M3& m) {
  a.l00 += m.r0.x; a.l10 += m.r1.x; a.l11 += m.r1.y; a.l20 += m.r2.x; a.l21 += m.r2.y; a.l22 += m.r2.z;
}
AVBD_HD void acc_ang_outer(Acc& a, V3 u, float s) {
  V3 su = u * s;
  a.a00 += su.x * u.x; a.a10 += su.y * u.x; a.a11 += su.y * u.y;
  a.a20 += su.z * u.x; a.a21 += su.z * u.y; a.a22 += su.z * u.z;
}
AVBD_HD void acc_ang_m3(Acc& a, const M3& m) {
  a.a00 += m.r0.x; a.a10 += m.r1.x; a.a11 += m.r1.y; a.a20 += m.r2.x; a.a21 += m.r2.y; a.a22 += m.r2.z;
}
AVBD_HD void acc_ang_diag3(Acc& a, V3 d) { a.a00 += d.x; a.a11 += d.y; a.a22 += d.z; }
// Hla += s * u v^T (u angular, v linear)
AVBD_HD void acc_cross_outer(Acc& a, V3 u, V3 v, float s) { a.cr = m3_add(a.cr, m3_scale(m3_outer(u, v), s)); }

AVBD_HD M3 lower_to_m3(float x00, float x10, float x11, float x20, float x21, float x22) {
  return M3{v3(x00, x10, x20), v3(x10, x11, x21), v3(x20, x21, x22)};
}

// Pose of a contact partner: the start-of-iteration snapshot when the
// manifold links two bodies of the same color (Jacobi coupling, flagged at
// adjacency build), else the live pose.  useSnap = false disables the rule
// (residual evaluation at a fixed iterate).
struct PoseView {
  const F4* lin;
  const F4* ang;
  const F4* linSnap;
  const F4* angSnap;
  bool useSnap;
  __device__ __forceinline__ F4 L(uint32_t b, uint32_t self, bool same) const {
    return (useSnap && same && b != self) ? linSnap[b] : lin[b];
  }
  __device__ __forceinline__ F4 A(uint32_t b, uint32_t self, bool same) const {
    return (useSnap && same && b != self) ? angSnap[b] : ang[b];
  }
};

// ---------------------------------------------------------------------------
// Stable Neo-Hookean tet (Smith et al. 2018) -- 40_solver.metal:972-1020.
// Psi = mu/2 (Ic-3) + lambda/2 (J - aT)^2, aT = 1 + mu/lambda; per-vertex
// Hessian  V(mu |w|^2 I + lambda g g^T)  with g = dJ/dx_self.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stamp_tet(const TetRec& t, uint32_t self, const F4* pos, Acc& acc) {
  V3 x0 = xyz(pos[t.ids.x]), x1 = xyz(pos[t.ids.y]), x2 = xyz(pos[t.ids.z]), x3 = xyz(pos[t.ids.w]);
  float muV = t.r1.w, lamV = t.r2.w;
  float alphaT = 1.0f + muV / fmaxf(lamV, 1e-9f);
  V3 d0 = x1 - x0, d1 = x2 - x0, d2 = x3 - x0;
  V3 m0 = xyz(t.r0), m1 = xyz(t.r1), m2 = xyz(t.r2);
  V3 f0 = d0 * m0.x + d1 * m1.x + d2 * m2.x;
  V3 f1 = d0 * m0.y + d1 * m1.y + d2 * m2.y;
  V3 f2 = d0 * m0.z + d1 * m1.z + d2 * m2.z;
  V3 j0 = cross(f1, f2), j1 = cross(f2, f0), j2 = cross(f0, f1);
  float J = dot(f0, j0);
  float s2 = lamV * (J - alphaT);
  V3 p0 = f0 * muV + j0 * s2, p1 = f1 * muV + j1 * s2, p2 = f2 * muV + j2 * s2;
  V3 w;
  if (self == t.ids.y) w = v3(m0.x, m0.y, m0.z);
  else if (self == t.ids.z) w = v3(m1.x, m1.y, m1.z);
  else if (self == t.ids.w) w = v3(m2.x, m2.y, m2.z);
  else w = -(m0 + m1 + m2);
  acc.gl += p0 * w.x + p1 * w.y + p2 * w.z;
  V3 g = j0 * w.x + j1 * w.y + j2 * w.z;
  acc_lin_diag(acc, muV * dot(w, w));
  acc_lin_outer(acc, g, lamV);
}

// The same stamp from a stamp-major slot: one gather level.
// e1..e3: the edges from the slot's own vertex to the other three
__device__ __forceinline__ void stamp_tet_slot_e(const TetSlot& t, V3 e1, V3 e2, V3 e3, Acc& acc) {
  V3 b1 = xyz(t.b1), b2 = xyz(t.b2), b3 = xyz(t.b3);
  float muV = t.b1.w, lamV = t.b2.w, alphaT = t.b3.w;
  V3 f0 = e1 * b1.x + e2 * b2.x + e3 * b3.x;
  V3 f1 = e1 * b1.y + e2 * b2.y + e3 * b3.y;
  V3 f2 = e1 * b1.z + e2 * b2.z + e3 * b3.z;
  V3 j0 = cross(f1, f2), j1 = cross(f2, f0), j2 = cross(f0, f1);
  float J = dot(f0, j0);
  float s2 = lamV * (J - alphaT);
  V3 w = -(b1 + b2 + b3);
  V3 g = j0 * w.x + j1 * w.y + j2 * w.z;  // dJ/dx_self
  acc.gl += (f0 * w.x + f1 * w.y + f2 * w.z) * muV + g * s2;
  acc_lin_diag(acc, muV * dot(w, w));
  acc_lin_outer(acc, g, lamV);
}
// edge from a vertex at (xs + ls) to one at (x + l), high and low parts
// differenced separately: neighbouring f32 positions subtract exactly, so
// the edge keeps the resolution of the low parts
__device__ __forceinline__ V3 split_edge(V3 x, V3 l, V3 xs, V3 ls) { return (x - xs) + (l - ls); }
__device__ __forceinline__ void stamp_tet_slot_x(const TetSlot& t, V3 xs, V3 x1, V3 x2, V3 x3, Acc& acc) {
  stamp_tet_slot_e(t, x1 - xs, x2 - xs, x3 - xs, acc);
}
__device__ __forceinline__ void stamp_tet_slot(const TetSlot& t, V3 xs, V3 ls, const F4* pos, const F4* lo,
                                               Acc& acc) {
  stamp_tet_slot_e(t, split_edge(xyz(pos[t.other.x]), xyz(lo[t.other.x]), xs, ls),
                   split_edge(xyz(pos[t.other.y]), xyz(lo[t.other.y]), xs, ls),
                   split_edge(xyz(pos[t.other.z]), xyz(lo[t.other.z]), xs, ls), acc);
}

// ---------------------------------------------------------------------------
// rigid / particle contact manifold -- 40_solver.metal:1155-1235
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 contact_force(const ManifoldRec& m, int ci, V3 dqAL, V3 dqAA, V3 dqBL, V3 dqBA, V3 rAW,
                                            V3 rBW, float alpha, V3 n, V3 t1, V3 t2, V3& Cout, float& fs,
                                            float& bounds) {
  const ContactRec& c = m.c[ci];
  V3 C = xyz(c.C0) * (1.0f - alpha);
  C += v3(dot(n, dqAL), dot(t1, dqAL), dot(t2, dqAL));
  C -= v3(dot(n, dqBL), dot(t1, dqBL), dot(t2, dqBL));
  C += v3(dot(cross(rAW, n), dqAA), dot(cross(rAW, t1), dqAA), dot(cross(rAW, t2), dqAA));
  C -= v3(dot(cross(rBW, n), dqBA), dot(cross(rBW, t1), dqBA), dot(cross(rBW, t2), dqBA));
  V3 pen = xyz(c.penalty), lam = xyz(c.lambda);
  V3 F = pen * C + lam;
  F.x = fminf(F.x, 0.0f);
  bounds = fabsf(F.x) * m.basisT1.w;
  fs = sqrtf(F.y * F.y + F.z * F.z);
  if (fs > bounds && fs > 0.0f) {
    float dyn = fabsf(F.x) * m.basisN.w;
    F.y *= dyn / fs;
    F.z *= dyn / fs;
  }
  Cout = C;
  return F;
}

// loSelf: the low part of the stamping body's position (added to its own
// displacement from the step start)
// contacts first, first + stride, ... of the manifold (a lane's share when a
// body's lanes split one manifold)
__device__ __forceinline__ void stamp_manifold(const ManifoldRec& m, uint32_t self, bool rigidSelf, const PoseView& pv,
                                               const F4* initLin, const F4* initAng, float alpha, Acc& acc,
                                               V3 loSelf, uint32_t first = 0u, uint32_t stride = 1u) {
  uint32_t a = m.header.x, b = m.header.y, n = m.header.z;
  bool isA = self == a;
  bool rA_round = (m.header.w & 2u) != 0, rB_round = (m.header.w & 4u) != 0;
  bool same = (m.header.w & kManifoldSameColor) != 0;
  F4 pAL = pv.L(a, self, same), pBL = pv.L(b, self, same), qA = pv.A(a, self, same), qB = pv.A(b, self, same);
  V3 dqAL = xyz(pAL) - xyz(initLin[a]);
  V3 dqAA = q_sub(qA, initAng[a]);
  V3 dqBL = xyz(pBL) - xyz(initLin[b]);
  if (isA) dqAL += loSelf;
  else dqBL += loSelf;
  V3 dqBA = q_sub(qB, initAng[b]);
  V3 nrm = xyz(m.basisN), t1 = xyz(m.basisT1), t2 = cross(nrm, t1);
  float s = isA ? 1.0f : -1.0f;
  for (uint32_t i = first; i < n; i += stride) {
    V3 rAW = rA_round ? xyz(m.c[i].rA) : q_rotate(qA, xyz(m.c[i].rA));
    V3 rBW = rB_round ? xyz(m.c[i].rB) : q_rotate(qB, xyz(m.c[i].rB));
    V3 C;
    float fs, bnd;
    V3 F = contact_force(m, (int)i, dqAL, dqAA, dqBL, dqBA, rAW, rBW, alpha, nrm, t1, t2, C, fs, bnd);
    V3 K = xyz(m.c[i].penalty);
    V3 jl0 = nrm * s, jl1 = t1 * s, jl2 = t2 * s;
    acc_lin_outer(acc, jl0, K.x);
    acc_lin_outer(acc, jl1, K.y);
    acc_lin_outer(acc, jl2, K.z);
    acc.gl += jl0 * F.x + jl1 * F.y + jl2 * F.z;
    if (rigidSelf) {
      V3 rW = isA ? rAW : rBW;
      V3 ja0 = cross(rW, jl0), ja1 = cross(rW, jl1), ja2 = cross(rW, jl2);
      acc_ang_outer(acc, ja0, K.x);
      acc_ang_outer(acc, ja1, K.y);
      acc_ang_outer(acc, ja2, K.z);
      acc_cross_outer(acc, ja0, jl0, K.x);
      acc_cross_outer(acc, ja1, jl1, K.y);
      acc_cross_outer(acc, ja2, jl2, K.z);
      acc.ga += ja0 * F.x + ja1 * F.y + ja2 * F.z;
    }
  }
}

// ---------------------------------------------------------------------------
// joints -- 40_solver.metal:741-908 (ball-socket / weld / hinge with motors,
// limits and armature)
// ---------------------------------------------------------------------------
AVBD_HD M3 geom_stiff_ball_socket(int k, V3 v) {
  M3 m = m3_diag(v3s(-comp(v, k)));
  if (k == 0) { m.r0.x += v.x; m.r1.x += v.y; m.r2.x += v.z; }
  else if (k == 1) { m.r0.y += v.x; m.r1.y += v.y; m.r2.y += v.z; }
  else { m.r0.z += v.x; m.r1.z += v.y; m.r2.z += v.z; }
  return m;
}
AVBD_HD float wrap_pi(float a) { return a - 6.2831853f * floorf((a + 3.14159265f) / 6.2831853f); }
AVBD_HD float hinge_twist(F4 qAr, F4 qB, V3 axis) {
  F4 r = q_mul(q_inv(qAr), qB);
  if (r.w < 0.0f) { r.x = -r.x; r.y = -r.y; r.z = -r.z; r.w = -r.w; }
  return 2.0f * atan2f(dot(v3(r.x, r.y, r.z), axis), r.w);
}

__device__ __forceinline__ void stamp_joint(const JointRec& j, uint32_t self, const F4* posLin, const F4* posAng,
                                            const F4* initAng, float alpha, float dt, Acc& acc) {
  uint32_t a = j.header.x, b = j.header.y;
  bool isA = self == a;
  float torqueArm = j.C0Lin.w;
  V3 penLin = xyz(j.penaltyLin);
  if (dot(penLin, penLin) > 0.0f) {
    V3 pA = a == kWorld ? xyz(j.rA) : xform(xyz(posLin[a]), posAng[a], xyz(j.rA));
    V3 pB = xform(xyz(posLin[b]), posAng[b], xyz(j.rB));
    V3 C = pA - pB;
    if (j.header.w & 1u) C -= xyz(j.C0Lin) * alpha;
    V3 F = penLin * C + xyz(j.lambdaLin);
    float js = isA ? 1.0f : -1.0f;
    V3 rW = isA ? q_rotate(posAng[a], xyz(j.rA)) : q_rotate(posAng[b], xyz(j.rB));
    M3 jAng = m3_skew(-(rW * js));
    M3 jAngT = m3_transpose(jAng);
    M3 K = m3_diag(penLin);
    M3 jAngTk = m3_mulm(jAngT, K);
    acc_lin_m3(acc, K);
    acc_ang_m3(acc, m3_mulm(jAngTk, jAng));
    acc.cr = m3_add(acc.cr, m3_scale(jAngTk, js));
    V3 r = rW * js;
    M3 H = m3_add(m3_add(m3_scale(geom_stiff_ball_socket(0, r), F.x), m3_scale(geom_stiff_ball_socket(1, r), F.y)),
                  m3_scale(geom_stiff_ball_socket(2, r), F.z));
    acc_ang_m3(acc, m3_diagonalize(H));
    acc.gl += F * js;
    acc.ga += m3_mul(jAngT, F);
  }
  V3 penAng = xyz(j.penaltyAng);
  if (dot(penAng, penAng) > 0.0f) {
    F4 qA = a == kWorld ? F4{0, 0, 0, 1} : posAng[a];
    F4 qAr = q_mul(qA, j.restRel);
    if (j.hingeAxis.w != 0.0f) {
      V3 axis = xyz(j.hingeAxis);
      V3 aB = q_rotate(posAng[b], axis);
      V3 aA = q_rotate(qAr, axis);
      V3 C = cross(aA, aB) * torqueArm;
      if (j.header.w & 2u) C -= xyz(j.C0Ang) * alpha;
      V3 F = penAng * C + xyz(j.lambdaAng);
      float alignment = dot(aA, aB);
      M3 Jm = isA ? m3_add(m3_outer(aA, aB), m3_scale(m3_identity(), -alignment))
                  : m3_add(m3_scale(m3_identity(), alignment), m3_scale(m3_outer(aB, aA), -1.0f));
      Jm = m3_scale(Jm, torqueArm);
      M3 JT = m3_transpose(Jm);
      acc_ang_m3(acc, m3_mulm(m3_mulm(JT, m3_diag(penAng)), Jm));
      acc.ga += m3_mul(JT, F);
      if (j.motor.y > 0.0f || j.limits.x < j.limits.y) {
        float twist = hinge_twist(qAr, posAng[b], axis);
        float sm = isA ? -1.0f : 1.0f;
        if (j.dynamics.x > 0.0f) {
          float Ca = wrap_pi(twist - j.dynamics.y);
          float ka = j.dynamics.x / fmaxf(dt * dt, 1e-12f);
          acc_ang_outer(acc, aB, ka);
          acc.ga += aB * (ka * Ca * sm);
        }
        if (j.limits.x < j.limits.y) {
          float over = fmaxf(twist - j.limits.y, 0.0f) + fminf(twist - j.limits.x, 0.0f);
          if (over != 0.0f) {
            float kL = 4.0e4f;
            float FL = clampf(kL * over, -3000.0f, 3000.0f);
            acc_ang_outer(acc, aB, kL);
            acc.ga += aB * (FL * sm);
          }
        }
        if (j.motor.y > 0.0f) {
          uint32_t mode = j.header.w & kMotorModeMask;
          if (mode == kMotorExplicitPD || mode == kMotorVelocity) {
            acc.ga += aB * (-j.dynamics.z * sm);
          } else {
            float Cm = wrap_pi(twist - j.motor.x);
            F4 qA0 = a == kWorld ? F4{0, 0, 0, 1} : initAng[a];
            float twist0 = hinge_twist(q_mul(qA0, j.restRel), initAng[b], axis);
            float dTwist = wrap_pi(twist - twist0);
            float cD = j.limits.z / fmaxf(dt, 1e-6f);
            float raw = j.motor.w * Cm + cD * dTwist;
            float effort = clampf(raw, -j.motor.y, j.motor.y);
            if (fabsf(raw) < j.motor.y) acc_ang_outer(acc, aB, j.motor.w + cD);
            acc.ga += aB * (effort * sm);
          }
        }
      }
    } else {
      V3 C = q_sub(qAr, posAng[b]) * torqueArm;
      float s = (isA ? 1.0f : -1.0f) * torqueArm;
      if (j.header.w & 2u) C -= xyz(j.C0Ang) * alpha;
      V3 F = penAng * C + xyz(j.lambdaAng);
      acc_ang_diag3(acc, penAng * (s * s));
      acc.ga += F * s;
    }
  }
}

// ---------------------------------------------------------------------------
// springs / hard rods -- 40_solver.metal:910-970
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stamp_spring(const SpringRec& sp, uint32_t self, bool rigidSelf, const F4* posLin,
                                             const F4* posAng, float alpha, Acc& acc) {
  uint32_t a = sp.header.x, b = sp.header.y;
  float stiffness = sp.rA.w, rest = sp.rB.w;
  V3 pA = xform(xyz(posLin[a]), posAng[a], xyz(sp.rA));
  V3 pB = xform(xyz(posLin[b]), posAng[b], xyz(sp.rB));
  V3 d = pA - pB;
  float dLen = length(d);
  if (dLen <= 1.0e-6f) return;
  V3 n = d / dLen;
  float f;
  if (sp.header.z != 0) {
    float C = dLen - rest - sp.dual.z * alpha;
    stiffness = sp.dual.y;
    f = stiffness * C + sp.dual.x;
    if (f <= 0.0f) return;
  } else {
    f = stiffness * (dLen - rest);
  }
  bool isA = self == a;
  V3 jLin = isA ? n : -n;
  acc_lin_outer(acc, jLin, stiffness);
  acc.gl += jLin * f;
  if (rigidSelf) {
    V3 rW = isA ? q_rotate(posAng[a], xyz(sp.rA)) : q_rotate(posAng[b], xyz(sp.rB));
    V3 jAng = isA ? cross(rW, n) : -cross(rW, n);
    acc_ang_outer(acc, jAng, stiffness);
    acc_cross_outer(acc, jAng, jLin, stiffness);
    acc.ga += jAng * f;
  }
  float ft = fmaxf(f, 0.0f) / dLen;
  if (ft > 0.0f) {
    acc_lin_diag(acc, ft);
    acc_lin_outer(acc, n, -ft);
  }
}

// ---------------------------------------------------------------------------
// cloth elements from stamp-major slots (40_solver.metal:910-970, 1088-1150)
// ---------------------------------------------------------------------------
// Partner position of a cloth slot: kSnapBit marks a partner of the same
// static color, read from the start-of-iteration snapshot.
__device__ __forceinline__ V3 slot_pos(const PoseView& pv, uint32_t id) {
  uint32_t b = id & ~kSnapBit;
  return xyz((pv.useSnap && (id & kSnapBit)) ? pv.linSnap[b] : pv.lin[b]);
}

// StVK membrane: Psi = mu A |E|^2 + lambda A / 2 tr(E)^2, E = (F^T F - I)/2,
// exact force P = F S; Gauss-Newton Hessian from the strain gradients
// (drops the indefinite geometric term).
__device__ __forceinline__ void stamp_membrane_slot(const ClothSlot& c, V3 xs, V3 x1, V3 x2, Acc& acc) {
  float muA = as_float(c.u.z), lamA = as_float(c.u.w);
  V3 e1 = x1 - xs, e2 = x2 - xs;
  V3 f1 = e1 * c.f.x + e2 * c.f.z;
  V3 f2 = e1 * c.f.y + e2 * c.f.w;
  float E11 = 0.5f * (dot(f1, f1) - 1.0f);
  float E22 = 0.5f * (dot(f2, f2) - 1.0f);
  float E12 = 0.5f * dot(f1, f2);
  float trE = E11 + E22;
  float S11 = 2.0f * muA * E11 + lamA * trE;
  float S22 = 2.0f * muA * E22 + lamA * trE;
  float S12 = 2.0f * muA * E12;
  V3 p1 = f1 * S11 + f2 * S12;
  V3 p2 = f1 * S12 + f2 * S22;
  float wx = -(c.f.x + c.f.z), wy = -(c.f.y + c.f.w);
  acc.gl += p1 * wx + p2 * wy;
  V3 g11 = f1 * wx, g22 = f2 * wy;
  V3 g12 = (f2 * wx + f1 * wy) * 0.5f;
  V3 gtr = g11 + g22;
  acc_lin_outer(acc, g11, 2.0f * muA);
  acc_lin_outer(acc, g22, 2.0f * muA);
  acc_lin_outer(acc, g12, 4.0f * muA);
  acc_lin_outer(acc, gtr, lamA);
}

// Quadratic (Bergou) bending: E = s/2 |sum_i K_i x_i|^2, evaluated
// translation-free as m = sum_o K_o (x_o - x_self) (K annihilates constants);
// grad_self = s K_self m, H_self = s K_self^2 I.
__device__ __forceinline__ void stamp_bend_slot(const ClothSlot& c, V3 xs, V3 x1, V3 x2, V3 x3, Acc& acc) {
  float s = as_float(c.u.w), ks = c.f.w;
  V3 m = (x1 - xs) * c.f.x + (x2 - xs) * c.f.y + (x3 - xs) * c.f.z;
  acc.gl += m * (s * ks);
  acc_lin_diag(acc, s * ks * ks);
}

// Particle-particle spring without offsets: soft k (|d| - L) or a
// tension-only augmented-Lagrangian rod, plus the geometric stiffness
// f/|d| (I - n n^T) while in tension.
__device__ __forceinline__ void stamp_pspring_slot(const ClothSlot& c, V3 xs, V3 xo, const SpringRec* springs,
                                                   float alpha, Acc& acc) {
  V3 d = xs - xo;
  float L = length(d);
  if (L <= 1.0e-6f) return;
  V3 n = d / L;
  float k = c.f.x, f;
  if (c.u.z != 0u) {
    F4 dual = springs[c.u.y].dual;
    float C = L - c.f.y - dual.z * alpha;
    k = dual.y;
    f = k * C + dual.x;
    if (f <= 0.0f) return;
  } else {
    f = k * (L - c.f.y);
  }
  acc_lin_outer(acc, n, k);
  acc.gl += n * f;
  float ft = fmaxf(f, 0.0f) / L;
  if (ft > 0.0f) {
    acc_lin_diag(acc, ft);
    acc_lin_outer(acc, n, -ft);
  }
}

__device__ __forceinline__ void stamp_cloth_slot(uint32_t kind, const ClothSlot& c, const PoseView& pv, V3 xs,
                                                 const SpringRec* springs, float alpha, Acc& acc) {
  if (kind == FK_BEND) {
    stamp_bend_slot(c, xs, slot_pos(pv, c.u.x), slot_pos(pv, c.u.y), slot_pos(pv, c.u.z), acc);
  } else if (kind == FK_MEMBRANE) {
    stamp_membrane_slot(c, xs, slot_pos(pv, c.u.x), slot_pos(pv, c.u.y), acc);
  } else {
    stamp_pspring_slot(c, xs, slot_pos(pv, c.u.x), springs, alpha, acc);
  }
}

// ---------------------------------------------------------------------------
// element contacts (OGC, 40_solver.metal:1237-1335): 4-slot weighted stencil
// ---------------------------------------------------------------------------
// `same` bit k: participant k shares the body's static color (snapshot pose).
__device__ __forceinline__ V3 soft_contact_C(const SoftRec& sc, const PoseView& pv, uint32_t same, const F4* initLin,
                                             const F4* initAng, float alpha, V3& n, V3& t1, V3& t2, V3& rAW) {
  uint32_t flags = as_uint(sc.anchorA.w);
  bool rigidA = (flags & kSoftRigidA) != 0, roundA = (flags & kSoftRoundA) != 0;
  n = xyz(sc.normal);
  orthonormal(n, t1, t2);
  rAW = v3s(0);
  V3 C = xyz(sc.C0) * (1.0f - alpha);
  auto add = [&](uint32_t b, float w, uint32_t k) {
    if (b == kWorld) return;
    bool snap = pv.useSnap && ((same >> k) & 1u);
    V3 dq = xyz(snap ? pv.linSnap[b] : pv.lin[b]) - xyz(initLin[b]);
    C += v3(dot(n, dq), dot(t1, dq), dot(t2, dq)) * w;
    if (k == 0 && rigidA) {
      F4 q = snap ? pv.angSnap[b] : pv.ang[b];
      rAW = roundA ? xyz(sc.anchorA) : q_rotate(q, xyz(sc.anchorA));
      V3 dqa = q_sub(q, initAng[b]);
      C += v3(dot(cross(rAW, n), dqa), dot(cross(rAW, t1), dqa), dot(cross(rAW, t2), dqa)) * w;
    }
  };
  add(sc.ids.x, sc.weights.x, 0);
  add(sc.ids.y, sc.weights.y, 1);
  add(sc.ids.z, sc.weights.z, 2);
  add(sc.ids.w, sc.weights.w, 3);
  return C;
}

// F = K C + lambda; below tau = restT / 2 the V-T / E-E normal switches to
// the log-barrier stage k' / d (k' = tau kc (r - tau), stitched C1);
// normal <= 0 and the Coulomb cone on the tangential pair.
__device__ __forceinline__ V3 soft_contact_force(const SoftRec& sc, V3 C, float& kNormal) {
  V3 F = xyz(sc.penalty) * C + xyz(sc.lambda);
  kNormal = sc.penalty.x;
  float rT = sc.anchorA.x;
  uint32_t kind = (as_uint(sc.anchorA.w) >> kSoftKindShift) & 3u;
  if (rT > 0.0f && kind != kSoftRT && C.x < -0.5f * rT) {
    float dsep = fmaxf(C.x + rT, 0.02f * rT);
    float kp = 0.25f * rT * rT * sc.penalty.x;
    F.x = -kp / dsep + sc.lambda.x;
    kNormal = kp / (dsep * dsep);
  }
  F.x = fminf(F.x, 0.0f);
  float bounds = fabsf(F.x) * sc.normal.w;
  float fs = sqrtf(F.y * F.y + F.z * F.z);
  if (fs > bounds && fs > 0.0f) {
    F.y *= bounds / fs;
    F.z *= bounds / fs;
  }
  return F;
}

__device__ __forceinline__ void stamp_soft(const SoftRec& sc, uint32_t self, uint32_t same, const PoseView& pv,
                                           const F4* initLin, const F4* initAng, float alpha, Acc& acc) {
  V3 n, t1, t2, rAW;
  V3 C = soft_contact_C(sc, pv, same, initLin, initAng, alpha, n, t1, t2, rAW);
  float kN;
  V3 F = soft_contact_force(sc, C, kN);
  float w;
  int slot;
  if (sc.ids.x == self) { slot = 0; w = sc.weights.x; }
  else if (sc.ids.y == self) { slot = 1; w = sc.weights.y; }
  else if (sc.ids.z == self) { slot = 2; w = sc.weights.z; }
  else if (sc.ids.w == self) { slot = 3; w = sc.weights.w; }
  else return;
  if (slot == 0 && (as_uint(sc.anchorA.w) & kSoftRigidA)) {
    V3 l0 = n * w, l1 = t1 * w, l2 = t2 * w;
    V3 a0 = cross(rAW, l0), a1 = cross(rAW, l1), a2 = cross(rAW, l2);
    float k1 = sc.penalty.y, k2 = sc.penalty.z;
    acc_lin_outer(acc, l0, kN);
    acc_lin_outer(acc, l1, k1);
    acc_lin_outer(acc, l2, k2);
    acc_ang_outer(acc, a0, kN);
    acc_ang_outer(acc, a1, k1);
    acc_ang_outer(acc, a2, k2);
    acc_cross_outer(acc, a0, l0, kN);
    acc_cross_outer(acc, a1, l1, k1);
    acc_cross_outer(acc, a2, l2, k2);
    // (angular first: both branches then end on the same gl update, so
    // the compiler cannot merge them into an address select on acc)
    acc.ga += a0 * F.x + a1 * F.y + a2 * F.z;
    acc.gl += l0 * F.x + l1 * F.y + l2 * F.z;
  } else {
    float w2 = w * w;
    acc_lin_outer(acc, n, kN * w2);
    acc_lin_outer(acc, t1, sc.penalty.y * w2);
    acc_lin_outer(acc, t2, sc.penalty.z * w2);
    acc.gl += (n * F.x + t1 * F.y + t2 * F.z) * w;
  }
}

// One adjacency entry; `slot` is the entry's static-CSR position (tets and
// cloth elements read their stamp-major record there).
// xs, ls: the stamping body's position, high and low parts
__device__ __forceinline__ void stamp_entry(const DeviceState& S, const PoseView& pv, uint32_t entry, uint32_t slot,
                                            uint32_t self, V3 xs, V3 ls, bool rigidSelf, float alpha, float dt,
                                            Acc& acc) {
  uint32_t kind = entry >> kAdjShift, idx = entry & kAdjIndexMask;
  if (kind == FK_TET) {
    stamp_tet_slot(S.tetSlots[slot], xs, ls, S.posLin, S.posLo, acc);
  } else if (kind == FK_MANIFOLD) {
    stamp_manifold(S.manifolds[idx], self, rigidSelf, pv, S.initLin, S.initAng, alpha, acc, ls);
  } else if (kind == FK_JOINT) {
    const JointRec& j = S.joints[idx];
    if (j.header.z == 0) stamp_joint(j, self, S.posLin, S.posAng, S.initAng, alpha, dt, acc);  // skip broken
  } else if (kind == FK_SPRING) {
    stamp_spring(S.springs[idx], self, rigidSelf, S.posLin, S.posAng, alpha, acc);
  } else if (kind == FK_SOFT) {
    stamp_soft(S.soft[entry & kSoftIndexMask], self, (entry >> kSoftSameShift) & 15u, pv, S.initLin, S.initAng, alpha,
               acc);
  } else if (kind == FK_MEMBRANE || kind == FK_BEND || kind == FK_PSPRING) {
    stamp_cloth_slot(kind, S.clothSlots[slot], pv, xs, S.springs, alpha, acc);
  }
}

}  // namespace avbd
