// util.hip -- fill and exclusive-scan primitives (10_scan.metal restated for
// wave64: 256 threads x 4 elements per block, block sums scanned by one
// block with a running carry, then added back).
#include "kernels.hpp"

namespace avbd {

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Several zero-fills in one launch (segment k covers grid-stride items).
struct FillSegs {
  uint32_t* p[6];
  uint32_t n[6];
};
__global__ void __launch_bounds__(256) k_zero_multi(FillSegs f) {
  for (int k = 0; k < 6; k++) {
    uint32_t* p = f.p[k];
    uint32_t n = f.n[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
  }
}

void launch_zero_multi(hipStream_t st, uint32_t* const* ptrs, const uint32_t* counts, int num) {
  FillSegs f{};
  uint32_t maxn = 0;
  for (int k = 0; k < num && k < 6; k++) {
    f.p[k] = ptrs[k];
    f.n[k] = counts[k];
    if (counts[k] > maxn) maxn = counts[k];
  }
  if (maxn == 0) return;
  uint32_t blocks = (maxn + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_zero_multi, dim3(blocks), dim3(256), 0, st, f);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t n, uint32_t value) {
  if (n == 0) return;
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, n, value);
}

constexpr uint32_t kScanTile = 1024;

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves):
// a wave64 shuffle scan, then the four wave totals combined through LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* lds, uint32_t& total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  // inclusive wave scan (Hillis-Steele over shuffles)
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  uint32_t waveOff = 0;
  for (uint32_t w = 0; w < wave; w++) waveOff += lds[w];
  total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return waveOff + x - v;
}

__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n,
                                                   uint32_t* blockSums) {
  __shared__ uint32_t lds[4];
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t a[4];
#pragma unroll
  for (int i = 0; i < 4; i++) a[i] = (base + i < n) ? in[base + i] : 0u;
  uint32_t local = a[0] + a[1] + a[2] + a[3];
  uint32_t total;
  uint32_t off = block_exclusive_scan_256(local, lds, total);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (base + i < n) out[base + i] = off;
    off += a[i];
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

// Scans the per-tile sums in place with one block, carrying across chunks.
__global__ void __launch_bounds__(256) k_scan_tile_sums(uint32_t* sums, uint32_t numTiles) {
  __shared__ uint32_t lds[4];
  uint32_t carry = 0;
  for (uint32_t chunk = 0; chunk < numTiles; chunk += 256) {
    uint32_t i = chunk + threadIdx.x;
    uint32_t v = i < numTiles ? sums[i] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan_256(v, lds, total);
    if (i < numTiles) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t* out, uint32_t n, const uint32_t* sums) {
  uint32_t add = sums[blockIdx.x];
  if (add == 0) return;
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (base + i < n) out[base + i] += add;
}

// Small inputs: one block of 1024 threads walks the array in tiles of 4096
// with a running carry -- one launch instead of three.
__global__ void __launch_bounds__(1024) k_scan_single(const uint32_t* in, uint32_t* out, uint32_t n) {
  __shared__ uint32_t waveSum[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += 4096) {
    uint32_t i0 = base + tid * 4;
    uint32_t a[4];
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = (i0 + i < n) ? in[i0 + i] : 0u;
    uint32_t local = a[0] + a[1] + a[2] + a[3];
    uint32_t x = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t y = __shfl_up(x, d, 64);
      if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) waveSum[wave] = x;
    __syncthreads();
    uint32_t off = carry, total = 0;
    for (uint32_t w = 0; w < 16; w++) {
      if (w < wave) off += waveSum[w];
      total += waveSum[w];
    }
    off += x - local;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (i0 + i < n) out[i0 + i] = off;
      off += a[i];
    }
    carry += total;
    __syncthreads();
  }
}

// Device-sized variants: the scanned length is min(*nDev, cap) + 1 (the
// extra element is a zero the producer wrote, so out[n - 1] is the total);
// tiles past it exit at once.
__device__ __forceinline__ uint32_t dev_len(const uint32_t* nDev, uint32_t cap) {
  uint32_t n = *nDev;
  return (n < cap ? n : cap) + 1u;
}
__global__ void __launch_bounds__(256) k_scan_tiles_dev(const uint32_t* in, uint32_t* out, const uint32_t* nDev,
                                                       uint32_t cap, uint32_t* blockSums) {
  __shared__ uint32_t lds[4];
  const uint32_t n = dev_len(nDev, cap);
  if (blockIdx.x * kScanTile >= n) return;
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
  uint32_t a[4];
#pragma unroll
  for (int i = 0; i < 4; i++) a[i] = (base + i < n) ? in[base + i] : 0u;
  uint32_t local = a[0] + a[1] + a[2] + a[3];
  uint32_t total;
  uint32_t off = block_exclusive_scan_256(local, lds, total);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (base + i < n) out[base + i] = off;
    off += a[i];
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}
__global__ void __launch_bounds__(256) k_scan_tile_sums_dev(uint32_t* sums, const uint32_t* nDev, uint32_t cap) {
  __shared__ uint32_t lds[4];
  const uint32_t n = dev_len(nDev, cap);
  const uint32_t numTiles = (n + kScanTile - 1) / kScanTile;
  uint32_t carry = 0;
  for (uint32_t chunk = 0; chunk < numTiles; chunk += 256) {
    uint32_t i = chunk + threadIdx.x;
    uint32_t v = i < numTiles ? sums[i] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan_256(v, lds, total);
    if (i < numTiles) sums[i] = carry + ex;
    carry += total;
  }
}
__global__ void __launch_bounds__(256) k_scan_add_dev(uint32_t* out, const uint32_t* nDev, uint32_t cap,
                                                     const uint32_t* sums) {
  const uint32_t n = dev_len(nDev, cap);
  if (blockIdx.x * kScanTile >= n) return;
  uint32_t add = sums[blockIdx.x];
  if (add == 0) return;
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (base + i < n) out[base + i] += add;
}
void launch_exclusive_scan_dev(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t nMax, const uint32_t* nDev,
                               uint32_t cap, uint32_t* scratch) {
  if (nMax == 0) return;
  uint32_t tiles = (nMax + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles_dev, dim3(tiles), dim3(256), 0, st, in, out, nDev, cap, scratch);
  hipLaunchKernelGGL(k_scan_tile_sums_dev, dim3(1), dim3(256), 0, st, scratch, nDev, cap);
  hipLaunchKernelGGL(k_scan_add_dev, dim3(tiles), dim3(256), 0, st, out, nDev, cap, scratch);
}

uint32_t scan_scratch_words(uint32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch) {
  if (n == 0) return;
  if (n <= 65536) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, in, out, n);
    return;
  }
  uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, st, in, out, n, scratch);
  if (tiles > 1) {
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(256), 0, st, scratch, tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(256), 0, st, out, n, scratch);
  }
}

}  // namespace avbd
