// Plain device-to-device copy of a known number of bytes: the reference point that scales the
// FETCH_SIZE / WRITE_SIZE hardware counters to HBM bytes when the benchmark is profiled.  Not
// part of the C-ABI of the BPNN path; loaded only by bench.py --copy-calibration.
#include <hip/hip_runtime.h>

namespace {

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

}  // namespace

// Copies `bytes` (rounded down to 16) from one fresh buffer to another `reps` times; 0 on success.
extern "C" int fnet_probe_copy(size_t bytes, int reps) {
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (n == 0 || reps <= 0) return 1;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return 2;
  if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
    (void)hipFree(a);
    return 2;
  }
  int rc = 0;
  if (hipMemset(a, 0, n * sizeof(float4)) != hipSuccess) rc = 3;
  for (int r = 0; r < reps && rc == 0; ++r) {
    k_copy<<<4096, 256>>>(a, b, n);
    if (hipGetLastError() != hipSuccess) rc = 4;
  }
  if (hipDeviceSynchronize() != hipSuccess) rc = 5;
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
