// predict.hip -- warm start of the step for gfx950: gamma warm start of the
// joint / rod duals and the inertial predict with the adaptive initial guess.
//
// Compiled without floating-point contraction (like the collision files):
// the predictor pose feeds the cloth pair stream, whose integer output must
// reproduce the oracle bit for bit.
#include "kernels.hpp"
#include "stamps.hpp"

namespace avbd {

// ---------------------------------------------------------------------------
// warm start (40_solver.metal:546-722)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_warmstart_constraints(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numJoints) {
    uint32_t si = gid - P.numJoints;
    if (si >= P.numSprings) return;
    SpringRec& sp = S.springs[si];
    if (sp.header.z == 0) return;
    uint32_t a = sp.header.x, b = sp.header.y;
    V3 pA = xform(xyz(S.posLin[a]), S.posAng[a], xyz(sp.rA));
    V3 pB = xform(xyz(S.posLin[b]), S.posAng[b], xyz(sp.rB));
    sp.dual.z = 0.5f * fmaxf(length(pA - pB) - sp.rB.w, 0.0f);
    float warm = P.alpha * P.gamma;
    sp.dual.x *= warm;
    sp.dual.y = fminf(clampf(sp.dual.y * P.gamma, kPenaltyMin, kPenaltyMax), sp.rA.w);
    return;
  }
  JointRec& j = S.joints[gid];
  if (j.header.z != 0) return;
  uint32_t a = j.header.x, b = j.header.y;
  V3 pA = a == kWorld ? xyz(j.rA) : xform(xyz(S.posLin[a]), S.posAng[a], xyz(j.rA));
  V3 pB = xform(xyz(S.posLin[b]), S.posAng[b], xyz(j.rB));
  F4 qA = a == kWorld ? F4{0, 0, 0, 1} : S.posAng[a];
  float torqueArm = j.C0Lin.w;
  j.C0Lin = f4(pA - pB, torqueArm);
  V3 c0a;
  V3 axis = xyz(j.hingeAxis);
  if (j.hingeAxis.w != 0.0f) {
    V3 aB = q_rotate(S.posAng[b], axis);
    V3 aA = q_rotate(q_mul(qA, j.restRel), axis);
    c0a = cross(aA, aB) * torqueArm;
    if (j.dynamics.x > 0.0f) {
      float twist = hinge_twist(q_mul(qA, j.restRel), S.posAng[b], axis);
      V3 omegaA = a == kWorld ? v3s(0) : xyz(S.velAng[a]);
      float rate = dot(xyz(S.velAng[b]) - omegaA, aB);
      j.dynamics.y = twist + P.dt * rate;
    }
  } else {
    c0a = q_sub(q_mul(qA, j.restRel), S.posAng[b]) * torqueArm;
  }
  j.C0Ang = f4(c0a, j.C0Ang.w);
  if (j.motor.y > 0.0f) {
    j.motor.z = 0.0f;
    j.dynamics.z = 0.0f;
    uint32_t mode = j.header.w & kMotorModeMask;
    if (mode == kMotorExplicitPD || mode == kMotorVelocity) {
      F4 qAr = a == kWorld ? j.restRel : q_mul(S.posAng[a], j.restRel);
      float twist = hinge_twist(qAr, S.posAng[b], axis);
      V3 axW = q_rotate(S.posAng[b], axis);
      V3 omegaA = a == kWorld ? v3s(0) : xyz(S.velAng[a]);
      float rate = dot(xyz(S.velAng[b]) - omegaA, axW);
      if (mode == kMotorVelocity)
        j.dynamics.z = clampf(j.limits.z * (j.motor.x - rate), -j.motor.y, j.motor.y);
      else
        j.dynamics.z = clampf(j.motor.w * (j.motor.x - twist) - j.limits.z * rate, -j.motor.y, j.motor.y);
    }
  }
  float warm = P.alpha * P.gamma;
  float stiffLin = j.rA.w, stiffAng = j.rB.w;
  j.lambdaLin.x *= warm; j.lambdaLin.y *= warm; j.lambdaLin.z *= warm;
  j.lambdaAng.x *= warm; j.lambdaAng.y *= warm; j.lambdaAng.z *= warm;
  j.penaltyLin.x = fminf(clampf(j.penaltyLin.x * P.gamma, kPenaltyMin, kPenaltyMax), stiffLin);
  j.penaltyLin.y = fminf(clampf(j.penaltyLin.y * P.gamma, kPenaltyMin, kPenaltyMax), stiffLin);
  j.penaltyLin.z = fminf(clampf(j.penaltyLin.z * P.gamma, kPenaltyMin, kPenaltyMax), stiffLin);
  j.penaltyAng.x = fminf(clampf(j.penaltyAng.x * P.gamma, kPenaltyMin, kPenaltyMax), stiffAng);
  j.penaltyAng.y = fminf(clampf(j.penaltyAng.y * P.gamma, kPenaltyMin, kPenaltyMax), stiffAng);
  j.penaltyAng.z = fminf(clampf(j.penaltyAng.z * P.gamma, kPenaltyMin, kPenaltyMax), stiffAng);
}

// Inertial predict x~ = x + v dt + g s dt^2, q~ = q (+) w dt; adaptive
// warm-start guess with the gravity weight clamp(a_z sgn(g)/|g|, 0, 1);
// start-of-step pose saved; also seeds the first iteration's snapshot.
__global__ void __launch_bounds__(256) k_warmstart_bodies(DeviceState S, StepParams P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= P.numBodies) return;
  F4 pl = S.posLin[gid], pa = S.posAng[gid];
  float mass = pl.w, dt = P.dt;
  V3 vl = xyz(S.velLin[gid]), va = xyz(S.velAng[gid]);
  V3 inertial = xyz(pl) + vl * dt;
  float g = P.gravity * S.gravityScale[gid];
  if (mass > 0.0f) inertial += v3(0, 0, g) * (dt * dt);
  S.inertLin[gid] = f4(inertial, 0.0f);
  F4 qi = q_addw(pa, va * dt);
  S.inertAng[gid] = qi;
  V3 accel = (vl - xyz(S.prevVelLin[gid])) / dt;
  float gsign = P.gravity < 0.0f ? -1.0f : (P.gravity > 0.0f ? 1.0f : 0.0f);
  float w = clampf(accel.z * gsign / fabsf(P.gravity), 0.0f, 1.0f);
  if (!finite_bits(w)) w = 0.0f;
  S.initLin[gid] = f4(xyz(pl), 0.0f);
  S.initAng[gid] = pa;
  // isotropic limiter: the step-start pose anchors the bounds of every
  // particle (pinned ones included), and the guess stays within the bound
  // (40_solver.metal:697-717)
  const bool isoPart = P.isoOn != 0u && S.shape[gid].w < 0.0f;
  if (isoPart) S.ogcPrev[gid] = f4(xyz(pl), 0.0f);
  F4 newL = pl, newA = pa;
  if (mass > 0.0f) {
    V3 guess = xyz(pl) + vl * dt + v3(0, 0, g) * (w * dt * dt);
    if (isoPart) {
      const float d2 = as_float(S.ogcBound[gid]);
      if (d2 < 1e37f) {
        const float bv = fmaxf(0.45f * sqrtf(d2), 0.2f * fabsf(S.shape[gid].w));
        const V3 tot = guess - xyz(pl);
        const float t2 = dot(tot, tot);
        if (t2 > bv * bv) guess = xyz(pl) + tot * (bv / sqrtf(t2));
      }
    }
    newL = f4(guess, mass);
    newA = qi;
    S.posLin[gid] = newL;
    S.posAng[gid] = newA;
  }
  S.posLo[gid] = F4{0.0f, 0.0f, 0.0f, 0.0f};
  S.posLinSnap[gid] = newL;
  S.posAngSnap[gid] = newA;
}

void launch_warmstart(hipStream_t st, const DeviceState& S, const StepParams& P) {
  uint32_t nc = P.numJoints + P.numSprings;
  if (nc > 0) hipLaunchKernelGGL(k_warmstart_constraints, dim3((nc + 255) / 256), dim3(256), 0, st, S, P);
  if (P.numBodies > 0)
    hipLaunchKernelGGL(k_warmstart_bodies, dim3((P.numBodies + 255) / 256), dim3(256), 0, st, S, P);
}

// Legacy servo rates (GPUSolver.swift:4407-4416): before the frame, each
// rate motor's target advances by rate * dt and is wrapped to [-pi, pi] by
// subtracting 2 pi times the nearest integer (ties away from zero); single
// roundings, no contraction, as the host arithmetic it restates.
__global__ void __launch_bounds__(256) k_rate_motors(DeviceState S, const uint32_t* joints, const float* rates,
                                                     uint32_t n, float dt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  JointRec& j = S.joints[joints[k]];
  const float twoPi = 6.28318530717958647692f;
  float target = __fadd_rn(j.motor.x, __fmul_rn(rates[k], dt));
  target = __fsub_rn(target, __fmul_rn(twoPi, roundf(__fdiv_rn(target, twoPi))));
  j.motor.x = target;
}

void launch_rate_motors(hipStream_t st, const DeviceState& S, const uint32_t* joints, const float* rates, uint32_t n,
                        float dt) {
  hipLaunchKernelGGL(k_rate_motors, dim3((n + 255) / 256), dim3(256), 0, st, S, joints, rates, n, dt);
}

}  // namespace avbd
