// HBM stream-mix probe: the bandwidth a kernel with the sweeps' access mix
// can reach when it does no stencil work.  Each mix reads NR and writes NW
// float arrays of the large slip-weakening grid's padded shape (x pitch
// 1088, 532 x 552 rows), one element per lane, marching 64 x 16 x-z tiles
// along y over 32 rows with the sweeps' XCD-aware block order, so the
// result is the ceiling of a sweep that moves the same words per cell.
// The flat_* mixes move the same words as 1-D float4 streams over the
// whole arrays (grid-stride, no tiles): the best rate those words reach.
// Array bases are skewed like the library's wavefield arrays.
// usage: hbm_stream [repetitions]
// Prints one JSON line per mix: bytes moved per launch, average ms, GB/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int NX = 1024, NY = 512, NZ = 512;   // cells marched (the interior of the grid)
constexpr int PX = 1088, NYT = 532, NZT = 552;  // array shape
constexpr int ROWS = 32;

__device__ inline int xcd_order(int b, int n) {
  const int q = n >> 3, r = n & 7, xcd = b & 7, loc = b >> 3;
  return xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

template <int NR, int NW>
__global__ void __launch_bounds__(1024) mix(float *const *in, float *const *out) {
  const int gx = NX / 64, gz = NZ / 16;
  const int b = xcd_order(blockIdx.x, gridDim.x);
  const int bx = b % gx, bz = (b / gx) % gz, by = b / (gx * gz);
  const int x = bx * 64 + threadIdx.x, z = bz * 16 + threadIdx.y;
  const long long sy = PX, sz = (long long)PX * NYT;
  for (int j = by * ROWS; j < by * ROWS + ROWS && j < NY; ++j) {
    const long long q = x + j * sy + z * sz;
    float v[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = in[r][q];
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < NR; ++r) s += v[r];
    if (NW == 0 && s == -1.2345e30f) in[0][q] = 0.0f;   // keeps read-only loads alive
#pragma unroll
    for (int w = 0; w < NW; ++w) __builtin_nontemporal_store(s * (float)(w + 1), &out[w][q]);
  }
}

// Flat form: every array as one 1-D stream of float4 lanes, grid-stride
// over the whole padded array, no tiles and no stencil shape: the rate the
// same words per cell reach at best.  NT: non-temporal stores.
typedef float f4 __attribute__((ext_vector_type(4)));
template <int NR, int NW, bool NT>
__global__ void __launch_bounds__(256) flat(f4 *const *in, f4 *const *out, long long n4) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n4; e += (long long)gridDim.x * 256) {
    f4 v[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = in[r][e];
    f4 s = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int r = 0; r < NR; ++r) s += v[r];
    if (NW == 0 && s.x == -1.2345e30f) in[0][e] = s;   // keeps read-only loads alive
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const f4 o = s * (float)(w + 1);
      if (NT) __builtin_nontemporal_store(o, &out[w][e]);
      else out[w][e] = o;
    }
  }
}

template <int NR, int NW, bool NT>
void run_flat(const char *name, std::vector<float *> &arr, size_t n, int reps) {
  std::vector<f4 *> in, out;
  for (int r = 0; r < NR; ++r) in.push_back(reinterpret_cast<f4 *>(arr[r]));
  for (int w = 0; w < NW; ++w) out.push_back(reinterpret_cast<f4 *>(arr[NR + w]));
  f4 **din = nullptr, **dout = nullptr;
  (void)hipMalloc(&din, sizeof(f4 *) * (NR + 1));
  (void)hipMalloc(&dout, sizeof(f4 *) * (NW + 1));
  if (NR) (void)hipMemcpy(din, in.data(), sizeof(f4 *) * NR, hipMemcpyHostToDevice);
  if (NW) (void)hipMemcpy(dout, out.data(), sizeof(f4 *) * NW, hipMemcpyHostToDevice);
  const long long n4 = (long long)(n / 4);
  const int blocks = 256 * 8;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL((flat<NR, NW, NT>), dim3(blocks), dim3(256), 0, 0, din, dout, n4);
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL((flat<NR, NW, NT>), dim3(blocks), dim3(256), 0, 0, din, dout, n4);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double bytes = 16.0 * (NR + NW) * (double)n4;
  printf("{\"mix\": \"%s\", \"read_words\": %d, \"written_words\": %d, \"bytes_per_launch\": %.0f, "
         "\"avg_ms\": %.4f, \"GBs\": %.1f}\n",
         name, NR, NW, bytes, ms / reps, bytes / (ms / reps * 1e-3) / 1e9);
  fflush(stdout);
  (void)hipFree(din);
  (void)hipFree(dout);
}

template <int NR, int NW>
void run(const char *name, std::vector<float *> &arr, int reps) {
  std::vector<float *> in(arr.begin(), arr.begin() + NR), out(arr.begin() + NR, arr.begin() + NR + NW);
  float **din = nullptr, **dout = nullptr;
  (void)hipMalloc(&din, sizeof(float *) * (NR + 1));
  (void)hipMalloc(&dout, sizeof(float *) * (NW + 1));
  if (NR) (void)hipMemcpy(din, in.data(), sizeof(float *) * NR, hipMemcpyHostToDevice);
  if (NW) (void)hipMemcpy(dout, out.data(), sizeof(float *) * NW, hipMemcpyHostToDevice);
  const int blocks = (NX / 64) * (NZ / 16) * ((NY + ROWS - 1) / ROWS);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL((mix<NR, NW>), dim3(blocks), dim3(64, 16), 0, 0, din, dout);
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL((mix<NR, NW>), dim3(blocks), dim3(64, 16), 0, 0, din, dout);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double bytes = 4.0 * (NR + NW) * (double)NX * NY * NZ;
  printf("{\"mix\": \"%s\", \"read_words\": %d, \"written_words\": %d, \"bytes_per_launch\": %.0f, "
         "\"avg_ms\": %.4f, \"GBs\": %.1f}\n",
         name, NR, NW, bytes, ms / reps, bytes / (ms / reps * 1e-3) / 1e9);
  fflush(stdout);
  (void)hipFree(din);
  (void)hipFree(dout);
}

}  // namespace

int main(int argc, char **argv) {
  const int reps = argc > 1 ? atoi(argv[1]) : 5;
  const size_t n = (size_t)PX * NYT * NZT;
  // array q starts (2304 q mod 65536) bytes into its allocation, as the
  // library skews its wavefield arrays (whole 256-byte units)
  std::vector<float *> arr(39), base(39);
  for (size_t q = 0; q < arr.size(); ++q) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(float) + 65536) != hipSuccess) {
      fprintf(stderr, "hipMalloc failed\n");
      return 1;
    }
    base[q] = static_cast<float *>(p);
    arr[q] = reinterpret_cast<float *>(static_cast<char *>(p) + (2304 * q) % 65536);
    (void)hipMemset(p, 0, n * sizeof(float) + 65536);
  }
  run<1, 1>("copy", arr, reps);
  run<8, 0>("read_only", arr, reps);
  run<9, 3>("velocity_sweep_mix", arr, reps);
  run<9, 6>("stress_sweep_mix", arr, reps);
  run<9, 9>("fused_sweep_mix", arr, reps);
  run<15, 12>("pml_velocity_mix", arr, reps);
  run<18, 21>("pml_stress_mix", arr, reps);
  // flat float4 streams over whole arrays
  run_flat<1, 1, true>("flat_copy", arr, n, reps);
  run_flat<1, 1, false>("flat_copy_cached_stores", arr, n, reps);
  run_flat<8, 0, true>("flat_read_only", arr, n, reps);
  run_flat<9, 3, true>("flat_velocity_sweep_mix", arr, n, reps);
  run_flat<9, 6, true>("flat_stress_sweep_mix", arr, n, reps);
  run_flat<9, 9, true>("flat_fused_sweep_mix", arr, n, reps);
  run_flat<9, 9, false>("flat_fused_sweep_mix_cached_stores", arr, n, reps);
  run_flat<15, 12, true>("flat_pml_velocity_mix", arr, n, reps);
  run_flat<18, 21, true>("flat_pml_stress_mix", arr, n, reps);
  for (auto p : base) (void)hipFree(p);
  return 0;
}
