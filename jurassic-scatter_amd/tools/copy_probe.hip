// copy_probe -- a plain device-to-device copy of known size, used to
// calibrate the HBM byte counters (FETCH_SIZE / WRITE_SIZE of rocprofv3)
// against the algorithmic traffic of the forward-model kernels.
//   copy_probe [MiB] [repetitions]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  size_t mib = argc > 1 ? strtoul(argv[1], nullptr, 10) : 1024;
  int reps = argc > 2 ? atoi(argv[2]) : 5;
  size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(k_copy, dim3(256 * 64), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  printf("copy_probe: %zu MiB read + %zu MiB written per launch, best %.3f ms, %.1f GB/s\n", mib, mib, best,
         2.0 * bytes / (best * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
