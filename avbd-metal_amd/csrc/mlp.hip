// mlp.hip -- the learned warm-start network on gfx950: per-row MLP
// 24 -> 128 -> 128 -> 3 or 6 (ReLU) with BF16 weights and activations and FP32
// accumulation on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16):
// hand-written forward, activation-gradient and split-K weight-gradient
// kernels, and Adam on the f32 master copy.
//
// The reference has no learned warm start; the network, its inputs and the
// fallback rule are this backend's own (avbd.h, learned.hip).
//
// MFMA 16x16x32 BF16 operand layouts (lane l of a wave64):
//   A (16 x 32, rows x k):  row l % 16, k = 8 (l / 16) + j, j = 0..7
//   B (32 x 16, k x cols):  col l % 16, k = 8 (l / 16) + j
//   D (16 x 16, f32):       col l % 16, row = 4 (l / 16) + i, i = 0..3
// so every operand fetch is one 16-byte load of 8 consecutive k.
//
// Rounding points (what a CPU restatement must reproduce): the inputs, the
// weights, the hidden activations h1 = bf16(relu(.)), h2 = bf16(relu(.)),
// and in the backward pass dY, dZ2 and dZ1 are rounded to bf16 (round to
// nearest even) before they feed a product; bias gradients use the f32 values.
#include "mlp.hpp"

namespace avbd {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kIn = kMlpIn, kInPad = kMlpInPad, kHid = kMlpHidden, kOutPad = kMlpOutPad;
// LDS row strides (bf16 elements): 80 B / 272 B / 48 B rows put the 16 rows
// of one operand fetch on 16 distinct 4-bank groups
constexpr int kS1 = 40, kSH = 136, kS3 = 24, kSD = 40;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ld8(const bf16* p) { return *(const bf16x8*)p; }
__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; j++) z[j] = (bf16)0.0f;
  return z;
}
__device__ __forceinline__ f32x4 zero4() { return f32x4{0.0f, 0.0f, 0.0f, 0.0f}; }
// LDS writes of one lane visible to the other lanes of its wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Weights as the kernels read them (bf16; global arrays dense, LDS copies padded):
//   w1t [kHid][kInPad]   layer-1 B operand (output-major, inputs contiguous)
//   w2t [kHid][kHid]     layer-2 B operand (output-major)
//   w3t [kOutPad][kHid]  layer-3 B operand (output-major)
//   w2  [kHid][kHid]     backward B operand of dH1 = dZ2 W2^T (input-major)
//   w3  [kHid][kOutPad]  backward B operand of dH2 = dY W3^T (input-major)
struct LdsFwd {
  bf16 w1t[kHid * kS1];
  bf16 w2t[kHid * kSH];
  bf16 w3t[kOutPad * kSH];
};
struct LdsBack {
  bf16 w2[kHid * kSH];
  bf16 w3[kHid * kS3];
};

// copy a dense [rows][cols] bf16 matrix into LDS rows of `stride` (16-byte chunks)
__device__ void stage(bf16* dst, const bf16* src, int rows, int cols, int stride) {
  const int chunks = cols / 8;
  for (int i = threadIdx.x; i < rows * chunks; i += blockDim.x) {
    int r = i / chunks, c = i % chunks;
    *(bf16x8*)(dst + r * stride + 8 * c) = *(const bf16x8*)(src + r * cols + 8 * c);
  }
}

__device__ void load_fwd(const MlpDevice& M, LdsFwd& L) {
  stage(L.w1t, M.w1t, kHid, kInPad, kS1);
  stage(L.w2t, M.w2t, kHid, kHid, kSH);
  stage(L.w3t, M.w3t, kOutPad, kHid, kSH);
}
__device__ void load_back(const MlpDevice& M, LdsBack& L) {
  stage(L.w2, M.w2, kHid, kHid, kSH);
  stage(L.w3, M.w3, kHid, kOutPad, kS3);
}

// Forward of one 16-row tile by one wave.  h1, h2: the wave's [16][kSH] LDS
// rows, left holding the bf16 activations.  Returns the layer-3 accumulator
// (D layout) plus bias.
__device__ f32x4 tile_forward(const MlpDevice& M, const LdsFwd& W, const bf16* X, uint32_t rows, uint32_t r0,
                              bf16* h1, bf16* h2, uint32_t lane) {
  const uint32_t ar = lane & 15u, ak = 8u * (lane >> 4), dr = 4u * (lane >> 4);
  bf16x8 a1 = (r0 + ar < rows) ? ld8(X + (size_t)(r0 + ar) * kInPad + ak) : zero8();
#pragma unroll
  for (int nt = 0; nt < kHid / 16; nt++) {
    f32x4 acc = mfma(a1, ld8(W.w1t + (nt * 16 + ar) * kS1 + ak), zero4());
    const uint32_t col = nt * 16 + ar;
    const float b = M.b1[col];
#pragma unroll
    for (int i = 0; i < 4; i++) h1[(dr + i) * kSH + col] = (bf16)fmaxf(acc[i] + b, 0.0f);
  }
  wave_sync();
#pragma unroll
  for (int nt = 0; nt < kHid / 16; nt++) {
    f32x4 acc = zero4();
#pragma unroll
    for (int kt = 0; kt < kHid / 32; kt++)
      acc = mfma(ld8(h1 + ar * kSH + kt * 32 + ak), ld8(W.w2t + (nt * 16 + ar) * kSH + kt * 32 + ak), acc);
    const uint32_t col = nt * 16 + ar;
    const float b = M.b2[col];
#pragma unroll
    for (int i = 0; i < 4; i++) h2[(dr + i) * kSH + col] = (bf16)fmaxf(acc[i] + b, 0.0f);
  }
  wave_sync();
  f32x4 acc = zero4();
#pragma unroll
  for (int kt = 0; kt < kHid / 32; kt++)
    acc = mfma(ld8(h2 + ar * kSH + kt * 32 + ak), ld8(W.w3t + ar * kSH + kt * 32 + ak), acc);
  const float b3 = M.b3[ar];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i] += b3;
  return acc;
}

// 16 rows of a wave's LDS rows, feature-major into dst[f * P + r0 ..]:
// lane l writes features l, l + 64, ... as 32 contiguous bytes each
__device__ void store_transposed(const bf16* h, int stride, int features, bf16* dst, size_t P, uint32_t r0,
                                 uint32_t lane) {
  for (int f = lane; f < features; f += 64) {
    bf16x8 lo, hi;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      lo[j] = h[j * stride + f];
      hi[j] = h[(8 + j) * stride + f];
    }
    *(bf16x8*)(dst + (size_t)f * P + r0) = lo;
    *(bf16x8*)(dst + (size_t)f * P + r0 + 8) = hi;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// inference: Y[r][0..nOut) = MLP(X[r]), X bf16 [rows][32], Y f32 [rows][nOut];
// 8 waves share one LDS copy of the weights
// ---------------------------------------------------------------------------
constexpr int kFwdWaves = 8;
__global__ void __launch_bounds__(64 * kFwdWaves) k_mlp_forward(MlpDevice M, const bf16* X, uint32_t rows, float* Y) {
  __shared__ LdsFwd W;
  __shared__ bf16 hbuf[kFwdWaves][2][16 * kSH];
  load_fwd(M, W);
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t tiles = (rows + 15) / 16;
  for (uint32_t t = blockIdx.x * kFwdWaves + wv; t < tiles; t += gridDim.x * kFwdWaves) {
    const uint32_t r0 = t * 16;
    f32x4 y = tile_forward(M, W, X, rows, r0, hbuf[wv][0], hbuf[wv][1], lane);
    const uint32_t col = lane & 15u;
    if (col < M.nOut) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint32_t r = r0 + 4 * (lane >> 4) + i;
        if (r < rows) Y[(size_t)r * M.nOut + col] = y[i];
      }
    }
    wave_sync();  // the next tile overwrites h1 / h2
  }
}

// ---------------------------------------------------------------------------
// training, pass 1: forward + activation gradients of
// L = sum_r |y_r - t_r|^2 * lossScale over every 16-row tile of the padded
// batch (rows >= `rows` carry zero gradient); writes the feature-major
// operands of the weight-gradient GEMMs and adds the bias gradients and
// sum |y - t|^2 atomically
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_act_grad(MlpDevice M, MlpTrain T, const bf16* X, const float* Tgt,
                                                      uint32_t rows, float lossScale) {
  __shared__ LdsFwd W;
  __shared__ LdsBack B;
  __shared__ bf16 hbuf[4][2][16 * kSH];
  __shared__ bf16 dzbuf[4][16 * kSD];
  load_fwd(M, W);
  load_back(M, B);
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t ar = lane & 15u, ak = 8u * (lane >> 4), dr = 4u * (lane >> 4);
  bf16 *h1 = hbuf[wv][0], *h2 = hbuf[wv][1], *dz = dzbuf[wv];
  float db1[kHid / 16], db2[kHid / 16], db3 = 0.0f, loss = 0.0f;
#pragma unroll
  for (int i = 0; i < kHid / 16; i++) db1[i] = db2[i] = 0.0f;
  const size_t P = T.rowsPad;
  const uint32_t tiles = T.rowsPad / 16;
  for (uint32_t t = blockIdx.x * 4 + wv; t < tiles; t += gridDim.x * 4) {
    const uint32_t r0 = t * 16;
    f32x4 y = tile_forward(M, W, X, rows, r0, h1, h2, lane);
    // dY = 2 (y - t) lossScale (D layout; columns >= nOut and padded rows are zero)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      uint32_t r = r0 + dr + i;
      float g = 0.0f;
      if (r < rows && ar < M.nOut) {
        float e = y[i] - Tgt[(size_t)r * M.nOut + ar];
        loss += e * e;
        g = 2.0f * e * lossScale;
      }
      db3 += g;
      dz[(dr + i) * kSD + ar] = (bf16)g;
      dz[(dr + i) * kSD + 16 + ar] = (bf16)0.0f;
    }
    // feature-major copy of X: lane = 16 rows x 4 groups of 8 features
    {
      const uint32_t r = r0 + (lane & 15u), f0 = 8u * (lane >> 4);
      bf16x8 xv = r < rows ? ld8(X + (size_t)r * kInPad + f0) : zero8();
#pragma unroll
      for (int j = 0; j < 8; j++) T.xt[(size_t)(f0 + j) * P + r] = xv[j];
    }
    wave_sync();
    store_transposed(h1, kSH, kHid, T.h1t, P, r0, lane);
    store_transposed(h2, kSH, kHid, T.h2t, P, r0, lane);
    store_transposed(dz, kSD, kOutPad, T.dyt, P, r0, lane);
    wave_sync();  // the transposed reads of h2 precede its in-place update
    // dH2 = dY W3^T (k = outputs, zero beyond 16); dZ2 = dH2 (h2 > 0) replaces h2 in place
    const bf16x8 ady = ld8(dz + ar * kSD + ak);
#pragma unroll
    for (int nt = 0; nt < kHid / 16; nt++) {
      bf16x8 bw = (lane < 32) ? ld8(B.w3 + (nt * 16 + ar) * kS3 + ak) : zero8();
      f32x4 acc = mfma(ady, bw, zero4());
      const uint32_t col = nt * 16 + ar;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t rr = dr + i;
        float g = ((float)h2[rr * kSH + col] > 0.0f && r0 + rr < rows) ? acc[i] : 0.0f;
        db2[nt] += g;
        h2[rr * kSH + col] = (bf16)g;  // each (row, col) is read and written by this lane only
      }
    }
    wave_sync();
    store_transposed(h2, kSH, kHid, T.dz2t, P, r0, lane);
    // dH1 = dZ2 W2^T; dZ1 = dH1 (h1 > 0), then replaces h1
#pragma unroll
    for (int nt = 0; nt < kHid / 16; nt++) {
      f32x4 acc = zero4();
#pragma unroll
      for (int kt = 0; kt < kHid / 32; kt++)
        acc = mfma(ld8(h2 + ar * kSH + kt * 32 + ak), ld8(B.w2 + (nt * 16 + ar) * kSH + kt * 32 + ak), acc);
      const uint32_t col = nt * 16 + ar;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t rr = dr + i;
        float g = ((float)h1[rr * kSH + col] > 0.0f && r0 + rr < rows) ? acc[i] : 0.0f;
        db1[nt] += g;
        h1[rr * kSH + col] = (bf16)g;  // h1 is not an operand of this product
      }
    }
    wave_sync();
    store_transposed(h1, kSH, kHid, T.dz1t, P, r0, lane);
    wave_sync();  // the next tile overwrites h1 / h2 / dz
  }
  // bias gradients: the four row groups of each column, then one atomic per column and wave
#pragma unroll
  for (int nt = 0; nt < kHid / 16; nt++) {
    float a = db1[nt], b = db2[nt];
    a += __shfl_xor(a, 16, 64);
    a += __shfl_xor(a, 32, 64);
    b += __shfl_xor(b, 16, 64);
    b += __shfl_xor(b, 32, 64);
    if (lane < 16) {
      atomicAdd(&T.gb1[nt * 16 + lane], a);
      atomicAdd(&T.gb2[nt * 16 + lane], b);
    }
  }
  db3 += __shfl_xor(db3, 16, 64);
  db3 += __shfl_xor(db3, 32, 64);
  if (lane < 16) atomicAdd(&T.gb3[lane], db3);
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
  if (lane == 0) atomicAdd(T.loss, loss);
}

// ---------------------------------------------------------------------------
// training, pass 2: weight gradients as split-K GEMMs over the samples,
//   gW1 = Xt dZ1t^T (32 x 128), gW2 = H1t dZ2t^T (128 x 128), gW3 = H2t dYt^T (128 x 16);
// 88 output tiles of 16 x 16, one per wave (blockIdx.y: 22 groups of 4),
// blockIdx.x owns a range of samples (a multiple of 32); partial sums added
// atomically (f32)
// ---------------------------------------------------------------------------
constexpr int kGradTiles = 2 * 8 + 8 * 8 + 8;
__global__ void __launch_bounds__(256) k_mlp_wgrad(MlpTrain T, uint32_t samplesPerBlock) {
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t ar = lane & 15u, ak = 8u * (lane >> 4), dr = 4u * (lane >> 4);
  const size_t P = T.rowsPad;
  const uint32_t k0 = blockIdx.x * samplesPerBlock;
  const uint32_t k1 = min((uint32_t)P, k0 + samplesPerBlock);
  const int tile = blockIdx.y * 4 + wv;
  if (tile >= kGradTiles) return;
  const bf16 *At, *Bt;
  float* G;
  uint32_t mt, nt, ldG;
  if (tile < 16) { At = T.xt; Bt = T.dz1t; G = T.gw1; mt = tile / 8; nt = tile % 8; ldG = kHid; }
  else if (tile < 80) { int q = tile - 16; At = T.h1t; Bt = T.dz2t; G = T.gw2; mt = q / 8; nt = q % 8; ldG = kHid; }
  else { int q = tile - 80; At = T.h2t; Bt = T.dyt; G = T.gw3; mt = q; nt = 0; ldG = kOutPad; }
  const bf16* ap = At + (size_t)(mt * 16 + ar) * P + ak;
  const bf16* bp = Bt + (size_t)(nt * 16 + ar) * P + ak;
  f32x4 acc = zero4();
  for (uint32_t k = k0; k < k1; k += 32) acc = mfma(ld8(ap + k), ld8(bp + k), acc);
#pragma unroll
  for (int i = 0; i < 4; i++) atomicAdd(&G[(mt * 16 + dr + i) * ldG + nt * 16 + ar], acc[i]);
}

// gradients in the master layout from the GEMM outputs (padding dropped)
__global__ void __launch_bounds__(256) k_mlp_gather_grad(MlpTrain T, uint32_t nOut) {
  const uint32_t offB3 = (uint32_t)mlp_off_b3((int)nOut), n = (uint32_t)mlp_params((int)nOut);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float g;
    if (i < (uint32_t)kMlpOffB1) g = T.gw1[i];  // rows 0..23 of gw1 [32][128]
    else if (i < (uint32_t)kMlpOffW2) g = T.gb1[i - kMlpOffB1];
    else if (i < (uint32_t)kMlpOffB2) g = T.gw2[i - kMlpOffW2];
    else if (i < (uint32_t)kMlpOffW3) g = T.gb2[i - kMlpOffB2];
    else if (i < offB3) {
      uint32_t q = i - kMlpOffW3;
      g = T.gw3[(q / nOut) * kOutPad + q % nOut];
    } else {
      g = T.gb3[i - offB3];
    }
    T.grad[i] = g;
  }
}

__global__ void __launch_bounds__(256) k_mlp_zero(MlpTrain T) {
  constexpr uint32_t n1 = kInPad * kHid, n2 = kHid * kHid, n3 = kHid * kOutPad;
  constexpr uint32_t total = n1 + n2 + n3 + 2 * kHid + kOutPad + 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    if (i < n1) T.gw1[i] = 0.0f;
    else if (i < n1 + n2) T.gw2[i - n1] = 0.0f;
    else if (i < n1 + n2 + n3) T.gw3[i - n1 - n2] = 0.0f;
    else if (i < n1 + n2 + n3 + kHid) T.gb1[i - n1 - n2 - n3] = 0.0f;
    else if (i < n1 + n2 + n3 + 2 * kHid) T.gb2[i - n1 - n2 - n3 - kHid] = 0.0f;
    else if (i < total - 1) T.gb3[i - n1 - n2 - n3 - 2 * kHid] = 0.0f;
    else T.loss[0] = 0.0f;
  }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the f32 master parameters
__global__ void __launch_bounds__(256) k_mlp_adam(MlpDevice M, MlpTrain T, float lr, float b1c, float b2c) {
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M.nParams; i += gridDim.x * blockDim.x) {
    float g = T.grad[i];
    float m = beta1 * T.m[i] + (1.0f - beta1) * g;
    float v = beta2 * T.v[i] + (1.0f - beta2) * g * g;
    T.m[i] = m;
    T.v[i] = v;
    M.params[i] -= lr * (m / b1c) / (sqrtf(v / b2c) + eps);
  }
}

// bf16 operand layouts and f32 biases from the f32 master parameters
__global__ void __launch_bounds__(256) k_mlp_pack(MlpDevice M) {
  const float* p = M.params;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)(kHid * kHid); i += gridDim.x * blockDim.x) {
    const uint32_t a = i / kHid, b = i % kHid;
    const float w = p[kMlpOffW2 + a * kHid + b];  // W2[in a][out b]
    M.w2[a * kHid + b] = (bf16)w;
    M.w2t[b * kHid + a] = (bf16)w;
    if (a < (uint32_t)kInPad) {
      const float w1 = a < (uint32_t)kIn ? p[kMlpOffW1 + a * kHid + b] : 0.0f;  // W1[in a][out b]
      M.w1t[b * kInPad + a] = (bf16)w1;
    }
    if (b < (uint32_t)kOutPad) {
      const float w3 = b < M.nOut ? p[kMlpOffW3 + a * M.nOut + b] : 0.0f;  // W3[in a][out b]
      M.w3[a * kOutPad + b] = (bf16)w3;
      M.w3t[b * kHid + a] = (bf16)w3;
    }
    if (a == 0) {
      M.b1[b] = p[kMlpOffB1 + b];
      M.b2[b] = p[kMlpOffB2 + b];
      if (b < (uint32_t)kOutPad) M.b3[b] = b < M.nOut ? p[mlp_off_b3((int)M.nOut) + b] : 0.0f;
    }
  }
}

// f32 rows [n][24] -> bf16 rows [n][32]
__global__ void __launch_bounds__(256) k_mlp_pack_input(const float* Xf, uint32_t n, bf16* X) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n * (uint32_t)kInPad; i += gridDim.x * blockDim.x) {
    const uint32_t r = i / kInPad, c = i % kInPad;
    X[i] = (bf16)(c < (uint32_t)kIn ? Xf[(size_t)r * kIn + c] : 0.0f);
  }
}

// batch rows by a permutation: four 16-byte chunks of X and the nOut targets per row
__global__ void __launch_bounds__(256) k_mlp_gather_batch(const bf16* Xall, const float* Tall, const uint32_t* perm,
                                                          uint32_t n, uint32_t nOut, bf16* X, float* T) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n; i += gridDim.x * blockDim.x) {
    const uint32_t r = i >> 2, c = i & 3u, src = perm[r];
    *(bf16x8*)(X + (size_t)r * kInPad + 8 * c) = *(const bf16x8*)(Xall + (size_t)src * kInPad + 8 * c);
    for (uint32_t k = c; k < nOut; k += 4) T[(size_t)r * nOut + k] = Tall[(size_t)src * nOut + k];
  }
}

// dst += loss * scale (one thread; per-epoch mean of the batch losses)
__global__ void k_mlp_accum_loss(const float* loss, float* dst, float scale) {
  if (threadIdx.x == 0) dst[0] += loss[0] * scale;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static uint32_t grid_for(uint32_t work, uint32_t cap) {
  uint32_t b = (work + 255) / 256;
  return b == 0 ? 1 : (b > cap ? cap : b);
}

void launch_mlp_pack(hipStream_t st, const MlpDevice& M) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, st, M);
}
void launch_mlp_pack_input(hipStream_t st, const float* Xf, uint32_t n, void* X) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_pack_input, dim3(grid_for(n * kInPad, 2048)), dim3(256), 0, st, Xf, n, (bf16*)X);
}
void launch_mlp_gather_batch(hipStream_t st, const void* Xall, const float* Tall, const uint32_t* perm, uint32_t n,
                             uint32_t nOut, void* X, float* T) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_gather_batch, dim3(grid_for(4 * n, 2048)), dim3(256), 0, st, (const bf16*)Xall, Tall,
                     perm, n, nOut, (bf16*)X, T);
}
void launch_mlp_forward(hipStream_t st, const MlpDevice& M, const void* X, uint32_t rows, float* Y, uint32_t cus) {
  if (rows == 0) return;
  const uint32_t tiles = (rows + 15) / 16;
  uint32_t blocks = (tiles + kFwdWaves - 1) / kFwdWaves;
  if (blocks > cus) blocks = cus;  // one 8-wave block per CU (LDS), tiles strided over the grid
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(64 * kFwdWaves), 0, st, M, (const bf16*)X, rows, Y);
}
void launch_mlp_grad(hipStream_t st, const MlpDevice& M, const MlpTrain& T, const void* X, const float* Tgt,
                     uint32_t rows, uint32_t cus) {
  hipLaunchKernelGGL(k_mlp_zero, dim3(96), dim3(256), 0, st, T);
  const uint32_t tiles = T.rowsPad / 16;
  uint32_t blocks = (tiles + 3) / 4;
  if (blocks > cus) blocks = cus;
  const float scale = 1.0f / ((float)M.nOut * (float)(rows > 0 ? rows : 1));
  hipLaunchKernelGGL(k_mlp_act_grad, dim3(blocks), dim3(256), 0, st, M, T, (const bf16*)X, Tgt, rows, scale);
  // split-K: at least 1024 samples per block, about one block per CU and tile group
  const uint32_t P = T.rowsPad;
  uint32_t spb = (P + cus - 1) / cus;
  spb = (spb + 31) / 32 * 32;
  if (spb < 1024) spb = 1024;
  const uint32_t kb = (P + spb - 1) / spb;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(kb, (kGradTiles + 3) / 4), dim3(256), 0, st, T, spb);
  hipLaunchKernelGGL(k_mlp_gather_grad, dim3(grid_for(M.nParams, 256)), dim3(256), 0, st, T, M.nOut);
}
void launch_mlp_accum_loss(hipStream_t st, const float* loss, float* dst, float scale) {
  hipLaunchKernelGGL(k_mlp_accum_loss, dim3(1), dim3(64), 0, st, loss, dst, scale);
}
void launch_mlp_adam(hipStream_t st, const MlpDevice& M, const MlpTrain& T, float lr, int step) {
  const float b1c = 1.0f - powf(0.9f, (float)step), b2c = 1.0f - powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3(grid_for(M.nParams, 256)), dim3(256), 0, st, M, T, lr, b1c, b2c);
  launch_mlp_pack(st, M);
}

}  // namespace avbd
