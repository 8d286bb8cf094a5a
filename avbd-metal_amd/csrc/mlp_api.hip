// mlp_api.hip -- C ABI of the learned warm-start network (avbd.h): create,
// parameters, forward, loss gradient, GPU training with Adam, forward timing.
// No reference counterpart.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/avbd.h"
#include "lifetime.hpp"
#include "mlp.hpp"

using namespace avbd;

namespace avbd {
void release_network(avbd_mlp* m);
}

namespace {

uint64_t splitmix(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform in [0, 1) from the top 24 bits
float unit01(uint64_t& s) { return (float)(splitmix(s) >> 40) * (1.0f / 16777216.0f); }

template <typename T>
hipError_t grab(std::vector<void*>& list, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(16, count * sizeof(T));
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return e;
  list.push_back(q);
  *p = (T*)q;
  return hipMemset(q, 0, bytes);
}

void release(std::vector<void*>& list) {
  for (void* p : list) hipFree(p);
  list.clear();
}

uint32_t round32(uint32_t n) { return (std::max<uint32_t>(n, 1u) + 31u) / 32u * 32u; }

// training buffers for batches of up to rowsPad rows
hipError_t ensure_train(avbd_mlp* m, uint32_t rowsPad) {
  if (m->T.rowsPad >= rowsPad && !m->trainAllocs.empty()) return hipSuccess;
  hipStreamSynchronize(m->stream);
  release(m->trainAllocs);
  MlpTrain T = m->T;
  const size_t P = rowsPad;
  hipError_t e;
  auto& L = m->trainAllocs;
  if ((e = grab(L, &T.xt, kMlpInPad * P)) || (e = grab(L, &T.h1t, kMlpHidden * P)) ||
      (e = grab(L, &T.h2t, kMlpHidden * P)) || (e = grab(L, &T.dz1t, kMlpHidden * P)) ||
      (e = grab(L, &T.dz2t, kMlpHidden * P)) || (e = grab(L, &T.dyt, kMlpOutPad * P)) ||
      (e = grab(L, &T.gw1, kMlpInPad * kMlpHidden)) || (e = grab(L, &T.gw2, kMlpHidden * kMlpHidden)) ||
      (e = grab(L, &T.gw3, kMlpHidden * kMlpOutPad)) || (e = grab(L, &T.gb1, kMlpHidden)) ||
      (e = grab(L, &T.gb2, kMlpHidden)) || (e = grab(L, &T.gb3, kMlpOutPad)) || (e = grab(L, &T.loss, 4)))
    return e;
  T.rowsPad = rowsPad;
  m->T = T;
  return hipSuccess;
}

// staging rows for forward / loss-gradient calls
hipError_t ensure_stage(avbd_mlp* m, size_t rows) {
  if (m->stageRows >= rows && !m->stageAllocs.empty()) return hipSuccess;
  hipStreamSynchronize(m->stream);
  release(m->stageAllocs);
  rows = std::max<size_t>(rows, 32);
  hipError_t e;
  __bf16* xb = nullptr;
  if ((e = grab(m->stageAllocs, &m->xf, rows * kMlpIn)) || (e = grab(m->stageAllocs, &xb, rows * kMlpInPad)) ||
      (e = grab(m->stageAllocs, &m->yf, rows * m->M.nOut)) || (e = grab(m->stageAllocs, &m->tf, rows * m->M.nOut)))
    return e;
  m->xb = xb;
  m->stageRows = rows;
  return hipSuccess;
}

avbd_status upload_rows(avbd_mlp* m, const float* x, uint32_t n) {
  if (ensure_stage(m, n) != hipSuccess) return AVBD_ERR_ALLOC;
  if (hipMemcpyAsync(m->xf, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return AVBD_ERR_HIP;
  launch_mlp_pack_input(m->stream, m->xf, n, m->xb);
  return AVBD_OK;
}

}  // namespace

extern "C" avbd_status avbd_mlp_create(uint64_t seed, int32_t outputs, avbd_mlp** out) {
  if (!out || (outputs != kMlpOut && outputs != kMlpOutRigid)) return AVBD_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return AVBD_ERR_NO_DEVICE;
  avbd_mlp* m = new avbd_mlp();
  auto fail = [&](avbd_status st) {
    release_network(m);
    return st;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail(AVBD_ERR_HIP);
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  m->cus = (uint32_t)std::max(cus, 1);
  MlpDevice& M = m->M;
  MlpTrain& T = m->T;
  M.nOut = (uint32_t)outputs;
  M.nParams = (uint32_t)mlp_params(outputs);
  const size_t np = M.nParams;
  auto& L = m->allocs;
  if (grab(L, &M.params, np) || grab(L, &M.w1t, kMlpHidden * kMlpInPad) ||
      grab(L, &M.w2t, kMlpHidden * kMlpHidden) || grab(L, &M.w3t, kMlpOutPad * kMlpHidden) ||
      grab(L, &M.w2, kMlpHidden * kMlpHidden) || grab(L, &M.w3, kMlpHidden * kMlpOutPad) ||
      grab(L, &M.b1, kMlpHidden) || grab(L, &M.b2, kMlpHidden) || grab(L, &M.b3, kMlpOutPad) ||
      grab(L, &T.grad, np) || grab(L, &T.m, np) || grab(L, &T.v, np))
    return fail(AVBD_ERR_ALLOC);
  // He-uniform weights (bound sqrt(6 / fan_in)); the output layer scaled by
  // 0.1 so an untrained network starts near the plain guess; zero biases
  std::vector<float> p(np, 0.0f);
  uint64_t s = seed;
  auto fill = [&](int off, int fanIn, int count, float scale) {
    const float a = scale * std::sqrt(6.0f / (float)fanIn);
    for (int i = 0; i < count; i++) p[off + i] = (2.0f * unit01(s) - 1.0f) * a;
  };
  fill(kMlpOffW1, kMlpIn, kMlpIn * kMlpHidden, 1.0f);
  fill(kMlpOffW2, kMlpHidden, kMlpHidden * kMlpHidden, 1.0f);
  fill(kMlpOffW3, kMlpHidden, kMlpHidden * outputs, 0.1f);
  avbd_status st = avbd_mlp_set_params(m, p.data());
  if (st != AVBD_OK) return fail(st);
  live_add(m, HandleKind::Network);
  *out = m;
  return AVBD_OK;
}

namespace avbd {

void release_network(avbd_mlp* m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  release(m->trainAllocs);
  release(m->stageAllocs);
  release(m->allocs);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

}  // namespace avbd

extern "C" void avbd_mlp_destroy(avbd_mlp* m) {
  // a handle already released at process exit is not touched again
  if (m && live_remove(m)) release_network(m);
}

extern "C" avbd_status avbd_mlp_info(const avbd_mlp* m, int32_t* outputs, int32_t* num_params) {
  if (!m) return AVBD_ERR_INVALID_ARGUMENT;
  if (outputs) *outputs = (int32_t)m->M.nOut;
  if (num_params) *num_params = (int32_t)m->M.nParams;
  return AVBD_OK;
}

extern "C" avbd_status avbd_mlp_get_params(avbd_mlp* m, float* params) {
  if (!m || !params) return AVBD_ERR_INVALID_ARGUMENT;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return AVBD_ERR_HIP;
  if (hipMemcpy(params, m->M.params, (size_t)m->M.nParams * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_mlp_set_params(avbd_mlp* m, const float* params) {
  if (!m || !params) return AVBD_ERR_INVALID_ARGUMENT;
  const size_t np = m->M.nParams;
  for (size_t i = 0; i < np; i++)
    if (!std::isfinite(params[i])) return AVBD_ERR_INVALID_ARGUMENT;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return AVBD_ERR_HIP;
  if (hipMemcpy(m->M.params, params, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->T.m, 0, np * 4) != hipSuccess || hipMemset(m->T.v, 0, np * 4) != hipSuccess)
    return AVBD_ERR_HIP;
  m->adamStep = 0;
  launch_mlp_pack(m->stream, m->M);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_mlp_forward(avbd_mlp* m, const float* x, uint32_t n, float* y) {
  if (!m || (n && (!x || !y))) return AVBD_ERR_INVALID_ARGUMENT;
  if (n == 0) return AVBD_OK;
  avbd_status st = upload_rows(m, x, n);
  if (st != AVBD_OK) return st;
  launch_mlp_forward(m->stream, m->M, m->xb, n, m->yf, m->cus);
  if (hipMemcpyAsync(y, m->yf, (size_t)n * m->M.nOut * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    return AVBD_ERR_HIP;
  return AVBD_OK;
}

extern "C" avbd_status avbd_mlp_loss_grad(avbd_mlp* m, const float* x, const float* t, uint32_t n, float* loss,
                                          float* grad) {
  if (!m || !x || !t || n == 0) return AVBD_ERR_INVALID_ARGUMENT;
  const uint32_t P = round32(n);
  if (ensure_train(m, P) != hipSuccess) return AVBD_ERR_ALLOC;
  avbd_status st = upload_rows(m, x, n);
  if (st != AVBD_OK) return st;
  if (hipMemcpyAsync(m->tf, t, (size_t)n * m->M.nOut * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return AVBD_ERR_HIP;
  MlpTrain T = m->T;
  T.rowsPad = P;
  launch_mlp_grad(m->stream, m->M, T, m->xb, m->tf, n, m->cus);
  float l = 0.0f;
  if (hipMemcpyAsync(&l, T.loss, 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
      (grad && hipMemcpyAsync(grad, T.grad, (size_t)m->M.nParams * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    return AVBD_ERR_HIP;
  if (loss) *loss = l / ((float)m->M.nOut * (float)n);
  return AVBD_OK;
}

extern "C" avbd_status avbd_mlp_train(avbd_mlp* m, const float* x, const float* t, uint32_t n, int32_t epochs,
                                      uint32_t batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!m || !x || !t || n == 0 || epochs <= 0 || batch == 0 || !(lr > 0.0f)) return AVBD_ERR_INVALID_ARGUMENT;
  const uint32_t B = std::min(batch, n);
  const uint32_t batches = std::max<uint32_t>(1u, n / B);
  const uint32_t P = round32(B);
  if (ensure_train(m, P) != hipSuccess) return AVBD_ERR_ALLOC;
  // the dataset (bf16 rows) and the targets stay resident for the whole run
  std::vector<void*> local;
  float *xf = nullptr, *tAll = nullptr, *tb = nullptr, *losses = nullptr;
  __bf16 *xAll = nullptr, *xbat = nullptr;
  uint32_t* perm = nullptr;
  auto done = [&](avbd_status st) {
    hipStreamSynchronize(m->stream);
    release(local);
    return st;
  };
  const uint32_t no = m->M.nOut;
  if (grab(local, &xf, (size_t)n * kMlpIn) || grab(local, &xAll, (size_t)n * kMlpInPad) ||
      grab(local, &tAll, (size_t)n * no) || grab(local, &xbat, (size_t)P * kMlpInPad) ||
      grab(local, &tb, (size_t)P * no) || grab(local, &losses, (size_t)epochs) ||
      grab(local, &perm, (size_t)epochs * n))
    return done(AVBD_ERR_ALLOC);
  // one shuffle per epoch (Fisher-Yates on SplitMix64), uploaded at once
  std::vector<uint32_t> order((size_t)epochs * n);
  uint64_t s = seed;
  for (int32_t e = 0; e < epochs; e++) {
    uint32_t* o = order.data() + (size_t)e * n;
    for (uint32_t i = 0; i < n; i++) o[i] = i;
    for (uint32_t i = n - 1; i > 0; i--) {
      uint32_t j = (uint32_t)(splitmix(s) % (uint64_t)(i + 1));
      std::swap(o[i], o[j]);
    }
  }
  if (hipMemcpyAsync(xf, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(tAll, t, (size_t)n * no * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(perm, order.data(), order.size() * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return done(AVBD_ERR_HIP);
  launch_mlp_pack_input(m->stream, xf, n, xAll);
  MlpTrain T = m->T;
  T.rowsPad = P;
  const float meanScale = 1.0f / ((float)no * (float)B * (float)batches);
  for (int32_t e = 0; e < epochs; e++) {
    for (uint32_t k = 0; k < batches; k++) {
      launch_mlp_gather_batch(m->stream, xAll, tAll, perm + (size_t)e * n + (size_t)k * B, B, no, xbat, tb);
      launch_mlp_grad(m->stream, m->M, T, xbat, tb, B, m->cus);
      launch_mlp_accum_loss(m->stream, T.loss, losses + e, meanScale);
      launch_mlp_adam(m->stream, m->M, T, lr, ++m->adamStep);
    }
  }
  if (hipGetLastError() != hipSuccess) return done(AVBD_ERR_HIP);
  if (epoch_loss &&
      hipMemcpyAsync(epoch_loss, losses, (size_t)epochs * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    return done(AVBD_ERR_HIP);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return done(AVBD_ERR_HIP);
  return done(AVBD_OK);
}

extern "C" avbd_status avbd_mlp_time_forward(avbd_mlp* m, uint32_t n, int32_t reps, float* avg_us) {
  if (!m || n == 0 || reps <= 0 || !avg_us) return AVBD_ERR_INVALID_ARGUMENT;
  std::vector<float> x((size_t)n * kMlpIn);
  uint64_t s = 7;
  for (float& v : x) v = 2.0f * unit01(s) - 1.0f;
  avbd_status st = upload_rows(m, x.data(), n);
  if (st != AVBD_OK) return st;
  launch_mlp_forward(m->stream, m->M, m->xb, n, m->yf, m->cus);  // warm
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return AVBD_ERR_HIP;
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return AVBD_ERR_HIP;
  }
  hipEventRecord(a, m->stream);
  for (int32_t r = 0; r < reps; r++) launch_mlp_forward(m->stream, m->M, m->xb, n, m->yf, m->cus);
  hipEventRecord(b, m->stream);
  hipError_t e = hipEventSynchronize(b);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (e != hipSuccess) return AVBD_ERR_HIP;
  *avg_us = ms * 1000.0f / (float)reps;
  return AVBD_OK;
}
