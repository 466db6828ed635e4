// bw_probe.hip -- STREAM-style copy and read kernels used by bench.py to
// measure the achievable HBM bandwidth of this GPU (reported beside the 8 TB/s
// spec) and to calibrate the FETCH_SIZE / WRITE_SIZE counters.
// Not part of the PolyDG path; built into liblymph_bwprobe.so.
#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 256;
constexpr int kVec = 4;  // double2 per thread, all loads issued before any use

__global__ __launch_bounds__(kThreads) void k_copy(const double2* __restrict__ a,
                                                   double2* __restrict__ b, size_t n) {
  const size_t base = (size_t)blockIdx.x * kThreads * kVec + threadIdx.x;
  double2 r[kVec];
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    if (i < n) r[k] = a[i];
  }
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    if (i < n) b[i] = r[k];
  }
}

__global__ __launch_bounds__(kThreads) void k_read(const double2* __restrict__ a,
                                                   double* __restrict__ out, size_t n) {
  const size_t base = (size_t)blockIdx.x * kThreads * kVec + threadIdx.x;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    if (i < n) {
      const double2 v = a[i];
      s += v.x + v.y;
    }
  }
  // one value per block keeps the loads live; lane 0 of each wave stores
  __shared__ double red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

// Read ceiling as the step kernel streams its blocks: non-temporal 16-byte
// loads, V per thread (8, 16 or 32), all issued before any use.
typedef double dv2_t __attribute__((ext_vector_type(2)));
template <int V>
__global__ __launch_bounds__(kThreads) void k_read_nt(const double2* __restrict__ a,
                                                      double* __restrict__ out, size_t n) {
  const size_t base = (size_t)blockIdx.x * kThreads * V + threadIdx.x;
  dv2_t v[V];
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    v[k] = i < n ? __builtin_nontemporal_load(reinterpret_cast<const dv2_t*>(a + i)) : dv2_t{0.0, 0.0};
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < V; ++k) s += v[k].x + v[k].y;
  __shared__ double red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

// Long-lived readers, as the step's waves are: `wgs` workgroups (a multiple
// of 8), each XCD's share walking a contiguous eighth of the buffer in
// chunks of kThreads x V double2, V non-temporal loads per thread in flight
// per chunk.
template <int V>
__global__ __launch_bounds__(kThreads) void k_read_persist(const double2* __restrict__ a,
                                                           double* __restrict__ out, size_t n) {
  const unsigned per_xcd = gridDim.x >> 3;
  const unsigned xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const size_t chunk = (size_t)kThreads * V;
  const size_t n_chunks = (n + chunk - 1) / chunk;
  const size_t c_per_xcd = (n_chunks + 7) / 8;
  const size_t c0 = xcd * c_per_xcd, c1 = c0 + c_per_xcd < n_chunks ? c0 + c_per_xcd : n_chunks;
  double s = 0.0;
  for (size_t c = c0 + w; c < c1; c += per_xcd) {
    const size_t base = c * chunk + threadIdx.x;
    dv2_t v[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const size_t i = base + (size_t)k * kThreads;
      v[k] = i < n ? __builtin_nontemporal_load(reinterpret_cast<const dv2_t*>(a + i)) : dv2_t{0.0, 0.0};
    }
#pragma unroll
    for (int k = 0; k < V; ++k) s += v[k].x + v[k].y;
  }
  __shared__ double red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w2 = 0; w2 < kThreads / 64; ++w2) t += red[w2];
    out[blockIdx.x] = t;
  }
}

constexpr unsigned kPersistWgs = 256 * 8;  // 8 workgroups of 4 waves per CU

// kind: 0 copy, 1 read, 8/16/32 non-temporal read with that many loads per
// thread, 116/132 the long-lived reader with 16/32 loads per thread
int run(int kind, size_t bytes, int reps, double* gbps, double* ms_per) {
  const bool copy = kind == 0;
  const bool persist = kind > 100;
  const int per = persist ? kind - 100 : (kind >= 8 ? kind : kVec);
  const size_t n = bytes / sizeof(double2);
  const size_t grid = persist ? kPersistWgs
                              : (n + (size_t)kThreads * per - 1) / ((size_t)kThreads * per);
  if (n == 0 || reps <= 0 || grid > 0x7fffffffu) return 1;
  double2* a = nullptr;
  double2* b = nullptr;
  double* out = nullptr;
  if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess) return 1;
  hipError_t e = copy ? hipMalloc(&b, n * sizeof(double2)) : hipMalloc(&out, grid * sizeof(double));
  if (e != hipSuccess) {
    (void)hipFree(a);
    return 1;
  }
  (void)hipMemset(a, 1, n * sizeof(double2));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&]() {
    if (copy)
      k_copy<<<(unsigned)grid, kThreads>>>(a, b, n);
    else if (kind == 1)
      k_read<<<(unsigned)grid, kThreads>>>(a, out, n);
    else if (kind == 8)
      k_read_nt<8><<<(unsigned)grid, kThreads>>>(a, out, n);
    else if (kind == 16)
      k_read_nt<16><<<(unsigned)grid, kThreads>>>(a, out, n);
    else if (kind == 32)
      k_read_nt<32><<<(unsigned)grid, kThreads>>>(a, out, n);
    else if (kind == 116)
      k_read_persist<16><<<(unsigned)grid, kThreads>>>(a, out, n);
    else
      k_read_persist<32><<<(unsigned)grid, kThreads>>>(a, out, n);
  };
  launch();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) launch();
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per = ms / reps;
  const double moved = (copy ? 2.0 : 1.0) * n * sizeof(double2);
  *gbps = moved / (*ms_per * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (out) (void)hipFree(out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // namespace

// Copy `bytes` (read + write counted) `reps` times; GB/s and ms per launch.
extern "C" int lymph_bw_copy(size_t bytes, int reps, double* gbps, double* ms_per) {
  return run(0, bytes, reps, gbps, ms_per);
}

// Read `bytes` `reps` times (one double written per 256-thread block).
extern "C" int lymph_bw_read(size_t bytes, int reps, double* gbps, double* ms_per) {
  return run(1, bytes, reps, gbps, ms_per);
}

// The same read with non-temporal 16-byte loads, 8, 16 and 32 in flight per
// thread, and by long-lived readers with 16 and 32; the fastest is reported.
extern "C" int lymph_bw_read_nt(size_t bytes, int reps, double* gbps, double* ms_per) {
  double best = 0.0, best_ms = 0.0;
  for (int v : {8, 16, 32, 116, 132}) {
    double g = 0.0, ms = 0.0;
    if (int rc = run(v, bytes, reps, &g, &ms)) return rc;
    if (g > best) {
      best = g;
      best_ms = ms;
    }
  }
  *gbps = best;
  *ms_per = best_ms;
  return 0;
}
