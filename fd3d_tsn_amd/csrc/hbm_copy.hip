// HBM copy probe: streams a float buffer of known size through a plain copy
// kernel.  Gives the achievable copy bandwidth of the device and, under
// rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, the counter reading of a known
// byte count (the calibration factor for the stencil kernels' traffic).
// usage: hbm_copy [MiB per buffer] [repetitions]
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_f4(const float4 *__restrict__ a, float4 *__restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char **argv) {
  long long mib = argc > 1 ? atoll(argv[1]) : 2048;
  int reps = argc > 2 ? atoi(argv[2]) : 20;
  long long bytes = mib << 20;
  long long n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  (void)hipMemset(a, 0, bytes);
  (void)hipMemset(b, 0, bytes);
  int blocks = 256 * 8 * 4;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  double gbs = 2.0 * bytes * reps / (ms * 1e-3) / 1e9;
  printf("{\"copy_bytes_per_launch\": %lld, \"read_bytes\": %lld, \"write_bytes\": %lld, "
         "\"avg_ms\": %.4f, \"GBs\": %.1f}\n", 2 * bytes, bytes, bytes, ms / reps, gbs);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
