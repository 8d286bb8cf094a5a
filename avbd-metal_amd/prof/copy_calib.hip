// copy_calib.hip -- calibration for rocprofv3 HBM counters: a plain device
// copy of a known size.  FETCH_SIZE / WRITE_SIZE read on this kernel give the
// bytes-per-count factor used to turn the same counters on the solver
// kernels into HBM bytes.
#include <hip/hip_runtime.h>
#include <stdio.h>

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int main() {
  const size_t bytes = 256ull << 20;
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  for (int r = 0; r < 3; r++) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  printf("copy_calib: 3 copies of %zu bytes\n", bytes);
  hipFree(a);
  hipFree(b);
  return 0;
}
