// upload.hip -- scene upload precompute on the GPU: the boundary faces of
// the tet meshes (GPUSolver.swift:872-897), the per-tet rest-shape inverse
// and the stamp-major slots of the tet entries of the static adjacency
// (GPUSolver.swift:1214-1236, 1476-1540).
//
// Boundary faces: every tet contributes its four faces as sorted vertex
// triples packed into one 63-bit key (21 bits per id); a device radix sort
// brings equal faces together, a face whose key occurs once lies on the
// boundary, it is wound away from its tet's fourth vertex, and a second
// sort of the oriented triples gives the ascending (a, b, c) order the
// collision triangles are listed in.  Scenes with 2^21 or more bodies keep
// the host path.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <array>
#include <vector>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace avbd {

namespace {

constexpr uint32_t kIdBits = 21;
constexpr uint64_t kIdMask = (1ull << kIdBits) - 1ull;

// faces of a tet (a, b, c) with the fourth vertex last
__constant__ uint8_t kFace[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};

__device__ __forceinline__ uint64_t pack3(uint32_t a, uint32_t b, uint32_t c) {
  return ((uint64_t)a << (2 * kIdBits)) | ((uint64_t)b << kIdBits) | (uint64_t)c;
}

__global__ void __launch_bounds__(256) k_face_keys(const uint32_t* ids, uint32_t nt, uint64_t* keys, uint32_t* vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4u * nt) return;
  const uint32_t t = i >> 2, f = i & 3u;
  uint32_t a = ids[4 * t + kFace[f][0]], b = ids[4 * t + kFace[f][1]], c = ids[4 * t + kFace[f][2]];
  // sort the three ids
  if (a > b) { uint32_t x = a; a = b; b = x; }
  if (b > c) { uint32_t x = b; b = c; c = x; }
  if (a > b) { uint32_t x = a; a = b; b = x; }
  keys[i] = pack3(a, b, c);
  vals[i] = i;
}

// 1 where the sorted key occurs exactly once
__global__ void __launch_bounds__(256) k_face_once(const uint64_t* keys, uint32_t n, uint32_t* flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const bool once = (i == 0 || keys[i - 1] != k) && (i + 1 == n || keys[i + 1] != k);
  flag[i] = once ? 1u : 0u;
}

// the boundary face in its tet's winding, turned to face away from the
// fourth vertex (same arithmetic as the host restatement: f32, no contraction)
__global__ void __launch_bounds__(256) k_face_orient(const uint32_t* ids, const float* pos, const uint32_t* vals,
                                                     const uint32_t* flag, const uint32_t* at, uint32_t n,
                                                     uint64_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || flag[i] == 0u) return;
  const uint32_t v = vals[i], t = v >> 2, f = v & 3u;
  uint32_t t0 = ids[4 * t + kFace[f][0]], t1 = ids[4 * t + kFace[f][1]], t2 = ids[4 * t + kFace[f][2]];
  const uint32_t opp = ids[4 * t + kFace[f][3]];
  auto P = [&](uint32_t b) { return v3(pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]); };
  const V3 pa = P(t0);
  const V3 nrm = cross(P(t1) - pa, P(t2) - pa);
  if (dot(nrm, P(opp) - pa) > 0.0f) { uint32_t x = t1; t1 = t2; t2 = x; }
  out[at[i]] = pack3(t0, t1, t2);
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)); }
};

}  // namespace

bool gpu_tet_boundary_faces(const uint32_t* tetIds, uint32_t nt, const float* bodyPos, uint32_t nb, hipStream_t st,
                            std::vector<std::array<uint32_t, 3>>& out) {
  out.clear();
  if (nt == 0) return true;
  if ((uint64_t)nb > kIdMask || 4ull * nt > 0x7FFFFFFFull) return false;
  const uint32_t n = 4u * nt;
  DevBuf<uint32_t> ids, vin, vout, flag, at;
  DevBuf<uint64_t> kin, kout, oriented, sorted;
  DevBuf<float> pos;
  DevBuf<char> tmp;
  if (ids.alloc(n) || pos.alloc(3 * (size_t)nb) || vin.alloc(n) || vout.alloc(n) || flag.alloc(n) ||
      at.alloc(n + 1) || kin.alloc(n) || kout.alloc(n))
    return false;
  if (hipMemcpyAsync(ids.p, tetIds, 4 * (size_t)n, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(pos.p, bodyPos, 12 * (size_t)nb, hipMemcpyHostToDevice, st))
    return false;
  const uint32_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_face_keys, dim3(blocks), dim3(256), 0, st, ids.p, nt, kin.p, vin.p);
  const int bits = 3 * kIdBits;
  size_t sortBytes = 0, scanBytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, sortBytes, kin.p, kout.p, vin.p, vout.p, n, 0, bits, st) ||
      rocprim::exclusive_scan(nullptr, scanBytes, flag.p, at.p, 0u, n + 0u, rocprim::plus<uint32_t>(), st))
    return false;
  if (tmp.alloc(std::max(sortBytes, scanBytes))) return false;
  if (rocprim::radix_sort_pairs(tmp.p, sortBytes, kin.p, kout.p, vin.p, vout.p, n, 0, bits, st)) return false;
  hipLaunchKernelGGL(k_face_once, dim3(blocks), dim3(256), 0, st, kout.p, n, flag.p);
  if (rocprim::exclusive_scan(tmp.p, scanBytes, flag.p, at.p, 0u, n + 0u, rocprim::plus<uint32_t>(), st))
    return false;
  uint32_t lastAt = 0, lastFlag = 0;
  if (hipMemcpyAsync(&lastAt, at.p + (n - 1), 4, hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(&lastFlag, flag.p + (n - 1), 4, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
    return false;
  const uint32_t nf = lastAt + lastFlag;
  if (nf == 0) return true;
  if (oriented.alloc(nf) || sorted.alloc(nf)) return false;
  hipLaunchKernelGGL(k_face_orient, dim3(blocks), dim3(256), 0, st, ids.p, pos.p, vout.p, flag.p, at.p, n,
                     oriented.p);
  size_t keyBytes = 0;
  if (rocprim::radix_sort_keys(nullptr, keyBytes, oriented.p, sorted.p, nf, 0, bits, st)) return false;
  DevBuf<char> tmp2;
  if (tmp2.alloc(keyBytes) || rocprim::radix_sort_keys(tmp2.p, keyBytes, oriented.p, sorted.p, nf, 0, bits, st))
    return false;
  std::vector<uint64_t> host(nf);
  if (hipMemcpyAsync(host.data(), sorted.p, 8 * (size_t)nf, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
    return false;
  out.resize(nf);
  for (uint32_t i = 0; i < nf; i++) {
    const uint64_t k = host[i];
    out[i] = {(uint32_t)(k >> (2 * kIdBits)), (uint32_t)((k >> kIdBits) & kIdMask), (uint32_t)(k & kIdMask)};
  }
  return true;
}

// Rest shape of every tet (GPUSolver.swift:1214-1236): rows of Dm^-1, the
// signed determinant (6 V), mu V and lambda V, in f64 and then rounded, as
// the host restatement computes them.
__global__ void __launch_bounds__(256) k_tet_rest(const uint32_t* ids, const float* pos, const float* mu,
                                                  const float* lam, uint32_t nt, TetRec* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const uint32_t* id = ids + 4 * t;
  double x0[3], D[3][3];
  for (int k = 0; k < 3; k++) x0[k] = pos[3 * id[0] + k];
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) D[k][c] = (double)pos[3 * id[c + 1] + k] - x0[k];
  const double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                     D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                     D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
  double inv[3][3];
  inv[0][0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) / det;
  inv[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det;
  inv[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det;
  inv[1][0] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) / det;
  inv[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det;
  inv[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det;
  inv[2][0] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) / det;
  inv[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det;
  inv[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det;
  const double vol = fabs(det) / 6.0;
  TetRec g;
  g.ids = U4{id[0], id[1], id[2], id[3]};
  g.r0 = F4{(float)inv[0][0], (float)inv[0][1], (float)inv[0][2], (float)det};
  g.r1 = F4{(float)inv[1][0], (float)inv[1][1], (float)inv[1][2], (float)(mu[t] * vol)};
  g.r2 = F4{(float)inv[2][0], (float)inv[2][1], (float)inv[2][2], (float)(lam[t] * vol)};
  out[t] = g;
}

// The stamp-major slot of every tet entry of the static adjacency: the
// other three vertices (in the tet's order) with their shape-function
// gradients (rows of Dm^-1; vertex 0 takes minus their sum), mu V, lambda V
// and 1 + mu / lambda.  One thread per body over its entries.
__global__ void __launch_bounds__(256) k_tet_slots(const TetRec* tets, const uint32_t* sStart, const uint32_t* sCount,
                                                   const uint32_t* sList, uint32_t nb, TetSlot* slots) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const uint32_t s0 = sStart[b], n = sCount[b];
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t e = sList[s0 + k];
    if ((e >> kAdjShift) != FK_TET) continue;
    const uint32_t t = e & kAdjIndexMask;
    const TetRec g = tets[t];
    const uint32_t ids[4] = {g.ids.x, g.ids.y, g.ids.z, g.ids.w};
    V3 grad[4];
    grad[1] = xyz(g.r0);
    grad[2] = xyz(g.r1);
    grad[3] = xyz(g.r2);
    grad[0] = -(grad[1] + grad[2] + grad[3]);
    uint32_t o[3] = {0u, 0u, 0u};
    V3 ob[3] = {v3s(0), v3s(0), v3s(0)};
    int m = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (ids[q] != b && m < 3) {
        if (m == 0) { o[0] = ids[q]; ob[0] = grad[q]; }
        else if (m == 1) { o[1] = ids[q]; ob[1] = grad[q]; }
        else { o[2] = ids[q]; ob[2] = grad[q]; }
        m++;
      }
    const float muV = g.r1.w, lamV = g.r2.w;
    TetSlot sl;
    sl.other = U4{o[0], o[1], o[2], t};
    sl.b1 = f4(ob[0], muV);
    sl.b2 = f4(ob[1], lamV);
    sl.b3 = f4(ob[2], 1.0f + muV / fmaxf(lamV, 1e-9f));
    slots[s0 + k] = sl;
  }
}

hipError_t gpu_tet_precompute(const uint32_t* tetIds, const float* mu, const float* lam, uint32_t nt,
                              const float* bodyPos, uint32_t nb, const uint32_t* sStartDev, const uint32_t* sCountDev,
                              const uint32_t* sListDev, TetRec* tetsDev, TetSlot* slotsDev, hipStream_t st) {
  if (nt == 0) return hipSuccess;
  DevBuf<uint32_t> ids;
  DevBuf<float> pos, m, l;
  hipError_t e;
  if ((e = ids.alloc(4 * (size_t)nt)) || (e = pos.alloc(3 * (size_t)nb)) || (e = m.alloc(nt)) || (e = l.alloc(nt)))
    return e;
  if ((e = hipMemcpyAsync(ids.p, tetIds, 16 * (size_t)nt, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(pos.p, bodyPos, 12 * (size_t)nb, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(m.p, mu, 4 * (size_t)nt, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(l.p, lam, 4 * (size_t)nt, hipMemcpyHostToDevice, st)))
    return e;
  hipLaunchKernelGGL(k_tet_rest, dim3((nt + 255) / 256), dim3(256), 0, st, ids.p, pos.p, m.p, l.p, nt, tetsDev);
  hipLaunchKernelGGL(k_tet_slots, dim3((nb + 255) / 256), dim3(256), 0, st, tetsDev, sStartDev, sCountDev, sListDev,
                     nb, slotsDev);
  if ((e = hipGetLastError())) return e;
  return hipStreamSynchronize(st);  // the staging buffers are freed on return
}

}  // namespace avbd
