// bw_probe.hip -- STREAM-style copy kernel used by bench.py to measure the
// achievable HBM bandwidth of the box (reported beside the 8 TB/s spec).
// Not part of the PolyDG path; built into liblymph_bwprobe.so.
#include <hip/hip_runtime.h>

__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ a,
                                              double2* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

extern "C" int lymph_bw_copy(size_t bytes, int reps, double* gbps, double* ms_per) {
  const size_t n = bytes / sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess) return 1;
  if (hipMalloc(&b, n * sizeof(double2)) != hipSuccess) { (void)hipFree(a); return 1; }
  (void)hipMemset(a, 1, n * sizeof(double2));
  (void)hipMemset(b, 0, n * sizeof(double2));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const unsigned grid = 256 * 8 * 4;
  k_copy<<<grid, 256>>>(a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) k_copy<<<grid, 256>>>(a, b, n);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per = ms / reps;
  *gbps = 2.0 * n * sizeof(double2) / (*ms_per * 1e-3) / 1e9;  // read + write
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
