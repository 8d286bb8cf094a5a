// mlp.hpp -- the learned warm start: a per-body MLP 24 -> 128 -> 128 -> 3
// (particles) or 6 (rigid bodies: linear + angular), ReLU, BF16 weights and
// activations, FP32 accumulation on the matrix cores, that predicts the
// converged correction of a step relative to the adaptive warm-start guess,
// in units of |g| dt^2 (angles as the displacement at the body's radius).
#pragma once
#include "state.hpp"

namespace avbd {

constexpr int kMlpIn = 24, kMlpHidden = 128, kMlpOut = 3, kMlpOutRigid = 6;
// f32 master parameters, layer-major: W1 [in][out], b1, W2 [in][out], b2,
// W3 [in][nOut], b3 [nOut]
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHidden;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
AVBD_HD constexpr int mlp_off_b3(int nOut) { return kMlpOffW3 + kMlpHidden * nOut; }
AVBD_HD constexpr int mlp_params(int nOut) { return mlp_off_b3(nOut) + nOut; }
constexpr int kMlpInPad = 32, kMlpOutPad = 16;

// Device copies of one network: the f32 master and the bf16 operand layouts
// the kernels read (see mlp.hip for each layout).
struct MlpDevice {
  uint32_t nOut, nParams;  // 3 or 6 outputs; mlp_params(nOut)
  float* params;
  __bf16 *w1t, *w2t, *w3t, *w2, *w3;
  float *b1, *b2, *b3;  // b3 zero-padded to 16
};

// Training buffers for batches of at most rowsPad rows (a multiple of 32):
// feature-major activations and gradients, GEMM outputs, Adam moments.
struct MlpTrain {
  __bf16 *xt, *h1t, *h2t, *dz1t, *dz2t, *dyt;
  float *gw1, *gw2, *gw3, *gb1, *gb2, *gb3, *loss;
  float *grad, *m, *v;
  uint32_t rowsPad;
};

// The learned stage inside a step: dynamic particles (and rigid bodies) in
// list order, their 1-ring (body ids; joint partners for rigid bodies), this
// step's plain guess and features, the previous step's converged correction
// per body, and the two start residuals.
struct WarmState {
  uint32_t* idx;          // [n] body id of list position i
  uint32_t *ringStart, *ringCount, *ringList;
  F4* guess;              // [n] plain adaptive guess (w = mass)
  F4* corr;               // [numBodies] last converged correction / L (xyz)
  F4* corrPrev;           // [numBodies] the one before (particles)
  __bf16* x;              // [n][32] network input rows
  float* y;               // [n][3] network output
  float* logX;            // [n][24] features (logging)
  float* logT;            // [n][3] targets (logging)
  float* res;             // [3] |g|^2 at the plain and the learned start, |m g_grav|^2
  F4 *pred, *pred2;       // [numBodies] predicted correction / L (ping-pong for the 1-ring smoothing)
  uint32_t n;
  uint32_t smoothPasses;  // Jacobi smoothing passes of the predicted correction over the 1-ring
  // rigid bodies
  uint32_t* ridx;         // [nR] body id of rigid list position i
  uint32_t *rRingStart, *rRingCount, *rRingList;  // joint partners
  F4 *rGuessLin, *rGuessAng;  // [nR] plain guess (position + mass, orientation)
  F4* corrAng;            // [numBodies] last angular correction (rotation vector x radius / L)
  __bf16* xR;             // [nR][32]
  float* yR;              // [nR][6]
  float *logXR, *logTR;   // [nR][24], [nR][6]
  uint32_t nR;
  float L;                // |g| dt^2 (feature and target unit)
  int force;              // 1: keep the learned start even when its residual is higher
  int log;                // 1: write logX / logT
};

void launch_mlp_pack(hipStream_t st, const MlpDevice& M);
void launch_mlp_pack_input(hipStream_t st, const float* Xf, uint32_t n, void* X);
void launch_mlp_forward(hipStream_t st, const MlpDevice& M, const void* X, uint32_t rows, float* Y, uint32_t cus);
// zeroes the gradient buffers, then forward + activation gradients + weight
// gradients of L = sum |y - t|^2 / (nOut rows) into T.grad (master layout) and T.loss
void launch_mlp_grad(hipStream_t st, const MlpDevice& M, const MlpTrain& T, const void* X, const float* Tgt,
                     uint32_t rows, uint32_t cus);
void launch_mlp_adam(hipStream_t st, const MlpDevice& M, const MlpTrain& T, float lr, int step);
void launch_mlp_accum_loss(hipStream_t st, const float* loss, float* dst, float scale);
// rows of a dataset gathered by a permutation: X bf16 [n][32], T f32 [n][nOut]
void launch_mlp_gather_batch(hipStream_t st, const void* Xall, const float* Tall, const uint32_t* perm,
                             uint32_t n, uint32_t nOut, void* X, float* T);

// learned stage of a step (learned.hip)
void launch_ws_features(hipStream_t st, const DeviceState& S, const StepParams& P, const WarmState& W);
// |g|^2 of the residual partials into W.res[slot]
void launch_ws_reduce(hipStream_t st, const DeviceState& S, uint32_t blocks, const WarmState& W, int slot);
// network output -> per-body correction, `smoothPasses` passes of
// c <- (c + <c_j>) / 2 over the 1-ring, then x_start = x_guess + L c
void launch_ws_apply(hipStream_t st, const DeviceState& S, const WarmState& W);
// rigid bodies: features, x_start = x_guess + L y_lin, q_start = q_guess (+) y_ang L / radius, targets
void launch_ws_features_rigid(hipStream_t st, const DeviceState& S, const StepParams& P, const WarmState& W);
void launch_ws_apply_rigid(hipStream_t st, const DeviceState& S, const WarmState& W);
void launch_ws_target_rigid(hipStream_t st, const DeviceState& S, const WarmState& W);
void launch_ws_select(hipStream_t st, const DeviceState& S, const WarmState& W);
// converged correction of this step (and the logged target)
void launch_ws_target(hipStream_t st, const DeviceState& S, const WarmState& W);

}  // namespace avbd

#include <vector>

// One network: its device copies, Adam state, training and staging buffers
// (grow-only), and its own stream (default priority).
struct avbd_mlp {
  avbd::MlpDevice M{};
  avbd::MlpTrain T{};        // T.rowsPad = capacity of the training buffers
  hipStream_t stream = nullptr;
  uint32_t cus = 256;
  int adamStep = 0;
  std::vector<void*> allocs;       // network + Adam buffers
  std::vector<void*> trainAllocs;  // feature-major training buffers
  std::vector<void*> stageAllocs;  // staging rows
  size_t stageRows = 0;
  float* xf = nullptr;  // [rows][24] f32
  void* xb = nullptr;   // [rows][32] bf16
  float* yf = nullptr;  // [rows][nOut]
  float* tf = nullptr;  // [rows][nOut]
};
