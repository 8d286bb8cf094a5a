// learned.hip -- the learned warm-start stage of a step.  After the
// adaptive guess (predict.hip) and the contact adjacency, each dynamic
// particle's features go through the MLP (mlp.hip), whose output replaces
// the guess: x_start = x_guess + L c with L = |g| dt^2 and c the network
// output after optional 1-ring smoothing passes.  The residual
// r = |g| / |m g_grav| is evaluated at both starts and the plain guess is
// kept whenever the learned start's residual is higher (or not finite).
// After the iterations, (x - x_guess) / L is the step's converged
// correction: the next step's feature and, when logging, the training target.
//
// Features (24, f32, in units of L; v_n, v_{n-1} velocities, x_n the
// step-start position, c and c' the previous two steps' corrections,
// <.> the mean over the topological 1-ring):
//   0-2  v_n dt / L               3-5   (v_n - v_{n-1}) dt / L
//   6-8  c'                       9-11  (x_guess - x_n) / L
//   12-14 c                       15-17 (<v_j> - v_n) dt / L
//   18-20 <c_j> - c               21    min(contact entries, 8) / 8
//   22   1-ring size / 16         23    |c|
// Rigid bodies (a 6-output network; r the body's bounding radius, angles as
// the displacement at r, c_lin / c_ang the previous corrections, <.> the
// mean over joint partners):
//   0-2  v_n dt / L               3-5   w_n dt r / L
//   6-8  (v_n - v_{n-1}) dt / L   9-11  (x_guess - x_n) / L
//   12-14 c_lin                   15-17 c_ang
//   18-20 <c_lin_j> - c_lin       21    min(contact entries, 8) / 8
//   22   min(joints, 4) / 4       23    |c_lin|
// x_start = x_guess + L y_lin, q_start = q_guess (+) y_ang L / r.
// The reference has no learned warm start; this stage is the backend's own.
#include <utility>

#include "kernels.hpp"
#include "mlp.hpp"

namespace avbd {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void __launch_bounds__(256) k_ws_features(DeviceState S, StepParams P, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.n) return;
  const uint32_t b = W.idx[i];
  const float invL = 1.0f / W.L, vs = P.dt * invL;
  const F4 g4 = S.posLin[b];
  const V3 x = xyz(S.initLin[b]), v = xyz(S.velLin[b]), c = xyz(W.corr[b]);
  V3 vSum = v3s(0.0f), cSum = v3s(0.0f);
  const uint32_t rs = W.ringStart[i], rc = W.ringCount[i];
  for (uint32_t k = 0; k < rc; k++) {
    const uint32_t j = W.ringList[rs + k];
    vSum += xyz(S.velLin[j]);
    cSum += xyz(W.corr[j]);
  }
  const float inv = rc ? 1.0f / (float)rc : 0.0f;
  const V3 f0 = v * vs;
  const V3 f1 = (v - xyz(S.prevVelLin[b])) * vs;
  const V3 f2 = xyz(W.corrPrev[b]);
  const V3 f3 = (xyz(g4) - x) * invL;
  const V3 f5 = rc ? (vSum * inv - v) * vs : v3s(0.0f);
  const V3 f6 = rc ? cSum * inv - c : v3s(0.0f);
  const float f[kMlpIn] = {f0.x, f0.y, f0.z, f1.x, f1.y, f1.z, f2.x, f2.y, f2.z, f3.x, f3.y, f3.z,
                           c.x,  c.y,  c.z,  f5.x, f5.y, f5.z, f6.x, f6.y, f6.z,
                           fminf((float)S.dAdjCount[b], 8.0f) * 0.125f, (float)rc * 0.0625f, length(c)};
  bf16x8 row[4];
#pragma unroll
  for (int k = 0; k < 32; k++) row[k / 8][k % 8] = (bf16)(k < kMlpIn ? f[k] : 0.0f);
#pragma unroll
  for (int q = 0; q < 4; q++) *(bf16x8*)(W.x + (size_t)i * kMlpInPad + 8 * q) = row[q];
  if (W.log) {
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) W.logX[(size_t)i * kMlpIn + k] = f[k];
  }
  W.guess[i] = g4;
}

// sums of the residual partials: |g|^2 into res[slot], |m g_grav|^2 into res[2]
__global__ void __launch_bounds__(256) k_ws_reduce(const double* partial, uint32_t blocks, float* res, int slot) {
  __shared__ double red[2][256];
  double g2 = 0.0, w2 = 0.0;
  for (uint32_t b = threadIdx.x; b < blocks; b += 256) {
    g2 += partial[2 * b];
    w2 += partial[2 * b + 1];
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
    res[slot] = (float)red[0][0];
    res[2] = (float)red[1][0];
  }
}

__global__ void __launch_bounds__(256) k_ws_pred(WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.n) return;
  W.pred[W.idx[i]] = F4{W.y[3 * i], W.y[3 * i + 1], W.y[3 * i + 2], 0.0f};
}

// one Jacobi pass c_i <- (c_i + <c_j>) / 2 over the 1-ring (bodies off the
// list keep a zero correction)
__global__ void __launch_bounds__(256) k_ws_smooth(WarmState W, const F4* in, F4* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.n) return;
  const uint32_t b = W.idx[i], rs = W.ringStart[i], rc = W.ringCount[i];
  const V3 c = xyz(in[b]);
  if (rc == 0) {
    out[b] = in[b];
    return;
  }
  V3 sum = v3s(0.0f);
  for (uint32_t k = 0; k < rc; k++) sum += xyz(in[W.ringList[rs + k]]);
  out[b] = f4((c + sum * (1.0f / (float)rc)) * 0.5f, 0.0f);
}

__global__ void __launch_bounds__(256) k_ws_apply(DeviceState S, WarmState W, const F4* pred) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.n) return;
  const uint32_t b = W.idx[i];
  const F4 g = W.guess[i];
  const F4 p = f4(xyz(g) + xyz(pred[b]) * W.L, g.w);
  S.posLin[b] = p;
  S.posLinSnap[b] = p;
}

// thread i restores particle i and rigid body i when the plain guess wins
__global__ void __launch_bounds__(256) k_ws_select(DeviceState S, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const float plain = W.res[0], learned = W.res[1];
  const bool keep = W.force || learned <= plain;  // false for a non-finite learned residual
  if (i == 0) {
    const float w2 = W.res[2] > 0.0f ? W.res[2] : 1.0f;
    S.counters[CTR_WS_ACCEPT] = keep ? 1u : 0u;
    S.counters[CTR_WS_RES_PLAIN] = as_uint(sqrtf(plain / w2));
    S.counters[CTR_WS_RES_LEARNED] = as_uint(sqrtf(learned / w2));
  }
  if (keep) return;
  if (i < W.n) {
    const uint32_t b = W.idx[i];
    S.posLin[b] = W.guess[i];
    S.posLinSnap[b] = W.guess[i];
  }
  if (i < W.nR) {
    const uint32_t b = W.ridx[i];
    S.posLin[b] = W.rGuessLin[i];
    S.posLinSnap[b] = W.rGuessLin[i];
    S.posAng[b] = W.rGuessAng[i];
    S.posAngSnap[b] = W.rGuessAng[i];
  }
}

__global__ void __launch_bounds__(256) k_ws_target(DeviceState S, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.n) return;
  const uint32_t b = W.idx[i];
  const V3 t = ((xyz(S.posLin[b]) - xyz(W.guess[i])) + xyz(S.posLo[b])) * (1.0f / W.L);
  W.corrPrev[b] = W.corr[b];
  W.corr[b] = f4(t, 0.0f);
  if (W.log) {
    W.logT[3 * i] = t.x;
    W.logT[3 * i + 1] = t.y;
    W.logT[3 * i + 2] = t.z;
  }
}

// ---------------------------------------------------------------------------
// rigid bodies
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rigid_radius(const DeviceState& S, uint32_t b) {
  return fmaxf(S.shape[b].w, 1.0e-3f);
}

__global__ void __launch_bounds__(256) k_ws_features_rigid(DeviceState S, StepParams P, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.nR) return;
  const uint32_t b = W.ridx[i];
  const float invL = 1.0f / W.L, vs = P.dt * invL, ws = P.dt * rigid_radius(S, b) * invL;
  const F4 g4 = S.posLin[b];
  const V3 x = xyz(S.initLin[b]), v = xyz(S.velLin[b]), w = xyz(S.velAng[b]);
  const V3 cl = xyz(W.corr[b]), ca = xyz(W.corrAng[b]);
  V3 cSum = v3s(0.0f);
  const uint32_t rs = W.rRingStart[i], rc = W.rRingCount[i];
  for (uint32_t k = 0; k < rc; k++) cSum += xyz(W.corr[W.rRingList[rs + k]]);
  const V3 f0 = v * vs;
  const V3 f1 = w * ws;
  const V3 f2 = (v - xyz(S.prevVelLin[b])) * vs;
  const V3 f3 = (xyz(g4) - x) * invL;
  const V3 f6 = rc ? cSum * (1.0f / (float)rc) - cl : v3s(0.0f);
  const float f[kMlpIn] = {f0.x, f0.y, f0.z, f1.x, f1.y, f1.z, f2.x, f2.y, f2.z, f3.x, f3.y, f3.z,
                           cl.x, cl.y, cl.z, ca.x, ca.y, ca.z, f6.x, f6.y, f6.z,
                           fminf((float)S.dAdjCount[b], 8.0f) * 0.125f, fminf((float)rc, 4.0f) * 0.25f, length(cl)};
  bf16x8 row[4];
#pragma unroll
  for (int k = 0; k < 32; k++) row[k / 8][k % 8] = (bf16)(k < kMlpIn ? f[k] : 0.0f);
#pragma unroll
  for (int q = 0; q < 4; q++) *(bf16x8*)(W.xR + (size_t)i * kMlpInPad + 8 * q) = row[q];
  if (W.log) {
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) W.logXR[(size_t)i * kMlpIn + k] = f[k];
  }
  W.rGuessLin[i] = g4;
  W.rGuessAng[i] = S.posAng[b];
}

__global__ void __launch_bounds__(256) k_ws_apply_rigid(DeviceState S, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.nR) return;
  const uint32_t b = W.ridx[i];
  const float* y = W.yR + (size_t)i * kMlpOutRigid;
  const F4 g = W.rGuessLin[i];
  const F4 p = f4(xyz(g) + v3(y[0], y[1], y[2]) * W.L, g.w);
  const F4 q = q_addw(W.rGuessAng[i], v3(y[3], y[4], y[5]) * (W.L / rigid_radius(S, b)));
  S.posLin[b] = p;
  S.posLinSnap[b] = p;
  S.posAng[b] = q;
  S.posAngSnap[b] = q;
}

__global__ void __launch_bounds__(256) k_ws_target_rigid(DeviceState S, WarmState W) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W.nR) return;
  const uint32_t b = W.ridx[i];
  const float invL = 1.0f / W.L;
  const V3 tl = ((xyz(S.posLin[b]) - xyz(W.rGuessLin[i])) + xyz(S.posLo[b])) * invL;
  const V3 ta = q_sub(S.posAng[b], W.rGuessAng[i]) * (rigid_radius(S, b) * invL);
  W.corr[b] = f4(tl, 0.0f);
  W.corrAng[b] = f4(ta, 0.0f);
  if (W.log) {
    float* t = W.logTR + (size_t)i * kMlpOutRigid;
    t[0] = tl.x;
    t[1] = tl.y;
    t[2] = tl.z;
    t[3] = ta.x;
    t[4] = ta.y;
    t[5] = ta.z;
  }
}

static dim3 grid_n(uint32_t n) { return dim3(n ? (n + 255) / 256 : 1); }

void launch_ws_features(hipStream_t st, const DeviceState& S, const StepParams& P, const WarmState& W) {
  if (W.n) hipLaunchKernelGGL(k_ws_features, grid_n(W.n), dim3(256), 0, st, S, P, W);
}
void launch_ws_reduce(hipStream_t st, const DeviceState& S, uint32_t blocks, const WarmState& W, int slot) {
  hipLaunchKernelGGL(k_ws_reduce, dim3(1), dim3(256), 0, st, (const double*)S.residualPartial, blocks, W.res, slot);
}
void launch_ws_apply(hipStream_t st, const DeviceState& S, const WarmState& W) {
  if (!W.n) return;
  hipLaunchKernelGGL(k_ws_pred, grid_n(W.n), dim3(256), 0, st, W);
  F4 *a = W.pred, *b = W.pred2;
  for (uint32_t k = 0; k < W.smoothPasses; k++) {
    hipLaunchKernelGGL(k_ws_smooth, grid_n(W.n), dim3(256), 0, st, W, (const F4*)a, b);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(k_ws_apply, grid_n(W.n), dim3(256), 0, st, S, W, (const F4*)a);
}
void launch_ws_select(hipStream_t st, const DeviceState& S, const WarmState& W) {
  hipLaunchKernelGGL(k_ws_select, grid_n(W.n > W.nR ? W.n : W.nR), dim3(256), 0, st, S, W);
}
void launch_ws_features_rigid(hipStream_t st, const DeviceState& S, const StepParams& P, const WarmState& W) {
  if (W.nR) hipLaunchKernelGGL(k_ws_features_rigid, grid_n(W.nR), dim3(256), 0, st, S, P, W);
}
void launch_ws_apply_rigid(hipStream_t st, const DeviceState& S, const WarmState& W) {
  if (W.nR) hipLaunchKernelGGL(k_ws_apply_rigid, grid_n(W.nR), dim3(256), 0, st, S, W);
}
void launch_ws_target_rigid(hipStream_t st, const DeviceState& S, const WarmState& W) {
  if (W.nR) hipLaunchKernelGGL(k_ws_target_rigid, grid_n(W.nR), dim3(256), 0, st, S, W);
}
void launch_ws_target(hipStream_t st, const DeviceState& S, const WarmState& W) {
  if (W.n) hipLaunchKernelGGL(k_ws_target, grid_n(W.n), dim3(256), 0, st, S, W);
}

}  // namespace avbd
