// mlp.hip -- liblymph_mlp.so: fully connected network (GELU) trained and
// evaluated with hand-written BF16 MFMA kernels for gfx950 (MI355X).
//
// GEMM: C[M][N] = sum_k A(m,k) B(k,n), bf16 operands, FP32 accumulation on
// v_mfma_f32_16x16x32_bf16.  TM x TM output tile (TM = 128, or 64 for small
// or split-k GEMMs) per 256-thread workgroup, each of the 4 waves owning a
// (TM/2)^2 quarter as (TM/32)^2 MFMA tiles; k advances 64 at a time through
// LDS tiles stored k-contiguous for both operands, so every lane's MFMA
// fragment is one 16-byte LDS read.  Operands may be k-contiguous
// in memory (row-major A, "W[n][k]" B: 16-byte loads copied as they are) or
// m/n-contiguous (16-byte loads along m/n, transposed into LDS), which covers
// the forward pass (X W^T), the input gradient (dZ W) and the weight
// gradient (dZ^T X) without transposed copies.  Epilogues fuse bias + GELU
// (forward, keeping the FP32 pre-activation) and the GELU derivative
// (backward).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lymph_mlp.h"

namespace {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

thread_local std::string g_err;
int fail(int code, std::string m) {
  g_err = std::move(m);
  return code;
}
#define HIP_TRY(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(LYMPH_MLP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// float -> bfloat16, round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

constexpr int BK = 64, LDK = BK + 8;  // LDS row: 64 k + 8 pad (144 B)
enum { EPI_F32 = 0, EPI_GELU_FWD = 1, EPI_GELU_BWD = 2 };

struct GemmArgs {
  int M, N, K;
  const bf16_t* A;  // A(m,k) = A_KC ? A[m*lda + k] : A[k*lda + m]
  int lda;
  const bf16_t* B;  // B(k,n) = B_KC ? B[n*ldb + k] : B[k*ldb + n]
  int ldb;
  float* C32;  // EPI_F32: result (+ bias);   EPI_GELU_FWD: pre-activation z
  bf16_t* C16; // EPI_GELU_FWD: gelu(z);      EPI_GELU_BWD: acc * gelu'(aux)
  int ldc;
  const float* bias;  // [N] or null
  const float* aux;   // EPI_GELU_BWD: pre-activation, ldc layout
  int kchunk;         // k range per blockIdx.z (multiple of BK)
  float* part;        // split-k partial sums [z][M][N], or null for a direct epilogue
};

// One TM x 64 operand tile in TM/32 16-byte pieces per thread: global ->
// registers (issued before the previous tile's MFMAs), registers -> LDS with
// k contiguous.
//  KC (k contiguous in memory): piece c = tid + 256 j is row c/8, k 8(c%8)..+7,
//    stored as it is.
//  not KC (rows contiguous in memory): thread tid owns rows 8 mg..8 mg+7 and
//    k rows RK kg..RK kg+RK-1 (mg = tid % (TM/8), kg = tid / (TM/8), RK = TM/32):
//    RK loads of 8 rows each, coalesced along the rows; in registers each row's
//    RK k values are gathered and written as one RK*2-byte LDS store.  The k
//    chunks (8 values) of row r are XOR-swizzled by (r/16) % 8 so that the 16
//    rows a wave writes at once fall in distinct banks; a 16-row MFMA tile has
//    one swizzle, so the fragment reads stay conflict-free.
template <bool KC>
__device__ inline int lds_k(int row, int k) {
  if (KC) return k;
  return ((((k >> 3) ^ (row >> 4)) & 7) << 3) | (k & 7);
}

template <bool KC, int TM>
__device__ inline void fetch(uint4 (&v)[TM / 32], const bf16_t* __restrict__ g, int ld, int r0,
                             int R, int k0, int ke, int tid) {
  constexpr int RK = TM / 32;
#pragma unroll
  for (int j = 0; j < RK; ++j) {
    v[j] = make_uint4(0u, 0u, 0u, 0u);
    if (KC) {
      const int c = tid + 256 * j;
      const int r = c >> 3, kc = (c & 7) * 8;
      if (r0 + r < R && k0 + kc < ke)
        v[j] = *reinterpret_cast<const uint4*>(g + (size_t)(r0 + r) * ld + k0 + kc);
    } else {
      const int mg = tid % (TM / 8), kk = (tid / (TM / 8)) * RK + j, rc = mg * 8;
      if (k0 + kk < ke && r0 + rc < R)
        v[j] = *reinterpret_cast<const uint4*>(g + (size_t)(k0 + kk) * ld + r0 + rc);
    }
  }
}

__device__ inline uint32_t word(const uint4& v, int w) {
  return w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
}

template <bool KC, int TM>
__device__ inline void put(bf16_t* __restrict__ s, const uint4 (&v)[TM / 32], int tid) {
  constexpr int RK = TM / 32;
  if (KC) {
#pragma unroll
    for (int j = 0; j < RK; ++j) {
      const int c = tid + 256 * j;
      *reinterpret_cast<uint4*>(s + (c >> 3) * LDK + (c & 7) * 8) = v[j];
    }
  } else {
    static_assert(RK == 2 || RK == 4, "tile rows");
    const int mg = tid % (TM / 8), kb = (tid / (TM / 8)) * RK;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = mg * 8 + q;
      // element q of k row j: half (q & 1) of word q / 2 of v[j]
      uint32_t h[RK];
#pragma unroll
      for (int j = 0; j < RK; ++j) h[j] = (q & 1) ? (word(v[j], q >> 1) >> 16) : (word(v[j], q >> 1) & 0xffffu);
      bf16_t* dst = s + row * LDK + lds_k<false>(row, kb);
      if constexpr (RK == 4)
        *reinterpret_cast<uint2*>(dst) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
      else
        *reinterpret_cast<uint32_t*>(dst) = h[0] | (h[1] << 16);
    }
  }
}

template <int EPI>
__device__ inline void epilogue(const GemmArgs& g, int m, int n, float v) {
  const size_t o = (size_t)m * g.ldc + n;
  if (EPI == EPI_F32) {
    g.C32[o] = g.bias ? v + g.bias[n] : v;
  } else if (EPI == EPI_GELU_FWD) {
    const float z = v + g.bias[n];
    g.C32[o] = z;
    g.C16[o] = f2bf(gelu(z));
  } else {
    g.C16[o] = f2bf(v * gelu_grad(g.aux[o]));
  }
}

// TM x TM output tile per 256-thread workgroup; 2 x 2 waves, each owning a
// (TM/2) x (TM/2) quarter as (TM/32) x (TM/32) MFMA tiles of 16 x 16.
template <bool A_KC, bool B_KC, int EPI, int TM>
__global__ __launch_bounds__(256, TM == 128 ? 3 : 1) void k_gemm(GemmArgs g) {
  constexpr int T = TM / 32;  // MFMA tiles per wave per dimension
  __shared__ __attribute__((aligned(16))) bf16_t sA[TM * LDK];
  __shared__ __attribute__((aligned(16))) bf16_t sB[TM * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TM;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TM / 2);
  const int kb = blockIdx.z * g.kchunk;
  const int ke = min(g.K, kb + g.kchunk);
  f32x4 acc[T][T];
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  uint4 ra[T], rb[T];
  fetch<A_KC, TM>(ra, g.A, g.lda, m0, g.M, kb, ke, tid);
  fetch<B_KC, TM>(rb, g.B, g.ldb, n0, g.N, kb, ke, tid);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    put<A_KC, TM>(sA, ra, tid);
    put<B_KC, TM>(sB, rb, tid);
    __syncthreads();
    if (k0 + BK < ke) {  // next tile in flight while this one is multiplied
      fetch<A_KC, TM>(ra, g.A, g.lda, m0, g.M, k0 + BK, ke, tid);
      fetch<B_KC, TM>(rb, g.B, g.ldb, n0, g.N, k0 + BK, ke, tid);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      bf16x8 af[T], bfg[T];
#pragma unroll
      for (int i = 0; i < T; ++i) {
        const int row = wm + i * 16 + fr;
        af[i] = *reinterpret_cast<const bf16x8*>(sA + row * LDK + lds_k<A_KC>(row, ks + fk));
      }
#pragma unroll
      for (int j = 0; j < T; ++j) {
        const int row = wn + j * 16 + fr;
        bfg[j] = *reinterpret_cast<const bf16x8*>(sB + row * LDK + lds_k<B_KC>(row, ks + fk));
      }
#pragma unroll
      for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfg[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D fragment: lane holds rows 4*(lane/16) + r, column lane % 16 of each 16x16 tile
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        if (g.part)
          g.part[((size_t)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][r];
        else
          epilogue<EPI>(g, m, n, acc[i][j][r]);
      }
}

// split-k: sum the S partial tiles in order, then the epilogue
template <int EPI>
__global__ void k_splitk(GemmArgs g, int S) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)g.M * g.N) return;
  const int m = (int)(i / g.N), n = (int)(i % g.N);
  float v = 0.f;
  for (int z = 0; z < S; ++z) v += g.part[((size_t)z * g.M + m) * g.N + n];
  epilogue<EPI>(g, m, n, v);
}

__global__ void k_gather(const float* __restrict__ X, int w, const int* __restrict__ idx, int B,
                         bf16_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)B * w) return;
  const int b = (int)(i / w), k = (int)(i % w);
  out[i] = f2bf(X[(size_t)idx[b] * w + k]);
}

// dY = scale * (y - t) in bf16; per-block partial sums of (y - t)^2.  Each
// thread takes 8 consecutive outputs of one row (N is a multiple of 8): two
// 16-byte loads of Y and of the target row, one 16-byte store of dY.
constexpr int kLossPer = 8 * 256;  // outputs per workgroup
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ Y, int N,
                                              const float* __restrict__ T,
                                              const int* __restrict__ idx, int B, float scale,
                                              bf16_t* __restrict__ dY, float* __restrict__ part) {
  __shared__ float red[256];
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
  float d2 = 0.f;
  if (i < (size_t)B * N) {
    const int b = (int)(i / N), n = (int)(i % N);
    const float4* y = reinterpret_cast<const float4*>(Y + i);
    const float4* t = reinterpret_cast<const float4*>(T + (size_t)idx[b] * N + n);
    const float4 y0 = y[0], y1 = y[1], t0 = t[0], t1 = t[1];
    const float d[8] = {y0.x - t0.x, y0.y - t0.y, y0.z - t0.z, y0.w - t0.w,
                        y1.x - t1.x, y1.y - t1.y, y1.z - t1.z, y1.w - t1.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      o[k] = (uint32_t)f2bf(scale * d[2 * k]) | ((uint32_t)f2bf(scale * d[2 * k + 1]) << 16);
      d2 += d[2 * k] * d[2 * k] + d[2 * k + 1] * d[2 * k + 1];
    }
    *reinterpret_cast<uint4*>(dY + i) = make_uint4(o[0], o[1], o[2], o[3]);
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// deterministic sum of n partials by one 256-thread workgroup: thread t adds
// partials t, t+256, ... in order, then a fixed tree over the threads
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ part, int n, float scale,
                                             float* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (float)(red[0] * scale);
}

// out[n] = sum_b D[b][n]: 64 columns per workgroup as 8 groups of 8 (one
// 16-byte load per row); 32 row groups per column group, row group g adding
// rows g, g + 32, ... in order, then the 32 partials of each column in order
// (deterministic).  N is a multiple of 8.
__global__ __launch_bounds__(256) void k_colsum(const bf16_t* __restrict__ D, int B, int N,
                                                float* __restrict__ out) {
  __shared__ float red[32][65];
  const int cg = threadIdx.x & 7, rg = threadIdx.x >> 3;
  const int n0 = blockIdx.x * 64 + cg * 8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (n0 < N)
    for (int b = rg; b < B; b += 32) {
      const uint4 v = *reinterpret_cast<const uint4*>(D + (size_t)b * N + n0);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += bf2f((bf16_t)(w[k] & 0xffffu));
        s[2 * k + 1] += bf2f((bf16_t)(w[k] >> 16));
      }
    }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[rg][cg * 8 + k] = s[k];
  __syncthreads();
  if (threadIdx.x < 64) {
    const int n = blockIdx.x * 64 + threadIdx.x;
    float t = 0.f;
    for (int g = 0; g < 32; ++g) t += red[g][threadIdx.x];
    if (n < N) out[n] = t;
  }
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ gr, bf16_t* __restrict__ wb, size_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ a, bf16_t* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct LymphMlp {
  int L = 0, max_batch = 0;
  std::vector<int> w;                  // widths
  std::vector<size_t> woff, boff;      // offsets of W_l, b_l in the parameter vector
  size_t n_params = 0;
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;  // master, grad, Adam
  bf16_t* Wb = nullptr;                // bf16 copy (W parts used; b parts unused)
  std::vector<bf16_t*> H;              // H[0] input .. H[L-1] (bf16 activations)
  std::vector<float*> Z;               // Z[0..L-2] pre-activations
  float* Y = nullptr;                  // [B][w_L]
  bf16_t *dY = nullptr, *dZa = nullptr, *dZb = nullptr;
  float *part = nullptr, *loss = nullptr;
  int* idx = nullptr;
  int n_data = 0;
  float *X = nullptr, *T = nullptr;
  int adam_t = 0;
  float* ws = nullptr;  // split-k partial sums, sized by the largest split GEMM seen
  size_t ws_floats = 0;
  // training's sample order and per-step losses, grown to the longest call
  int* dorder = nullptr;
  size_t dorder_cap = 0;
  float* dloss = nullptr;
  size_t dloss_cap = 0;
  std::vector<void*> owned;
  // every kernel and copy of this network runs on its own stream, so that
  // networks trained from different host threads overlap on the GPU
  hipStream_t stream = nullptr;
};

namespace {

template <class T>
int dalloc(LymphMlp* m, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess)
    return fail(LYMPH_MLP_ERR_HIP, "hipMalloc(" + std::to_string(n * sizeof(T)) + " B) failed");
  m->owned.push_back((void*)*p);
  return 0;
}

// Launch: 128 x 128 tiles when they still give >= 256 workgroups, else
// 64 x 64; a GEMM with few tiles and a long k range (the input gradient of the
// output layer: k = 38,400) is split along k over up to kMaxSplit workgroups
// per tile (32 x 32 = 1024 workgroups there) and reduced in order.
constexpr int kMaxSplit = 32;
template <bool A_KC, bool B_KC, int EPI>
int gemm(LymphMlp* m, GemmArgs a) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0) return 0;
  const bool big = blocks(a.N, 128) * blocks(a.M, 128) >= 256;
  const int TM = big ? 128 : 64;
  const unsigned tiles = blocks(a.N, TM) * blocks(a.M, TM);
  int S = 1;
  if (tiles < 256 && a.K > 1024) S = std::min<int>(kMaxSplit, (a.K + 1023) / 1024);
  a.kchunk = (int)(blocks(blocks(a.K, S), BK) * BK);
  S = (int)blocks(a.K, a.kchunk);
  a.part = nullptr;
  if (S > 1) {
    const size_t need = (size_t)S * a.M * a.N;
    if (need > m->ws_floats) {  // grown once, on the first call of a larger shape
      (void)hipStreamSynchronize(m->stream);
      if (m->ws) (void)hipFree(m->ws);
      m->ws = nullptr;
      m->ws_floats = 0;
      if (hipMalloc(&m->ws, need * sizeof(float)) != hipSuccess)
        return fail(LYMPH_MLP_ERR_HIP, "hipMalloc of the split-k workspace failed");
      m->ws_floats = need;
    }
    a.part = m->ws;
  }
  dim3 grid(blocks(a.N, TM), blocks(a.M, TM), S);
  if (big)
    k_gemm<A_KC, B_KC, EPI, 128><<<grid, 256, 0, m->stream>>>(a);
  else
    k_gemm<A_KC, B_KC, EPI, 64><<<grid, 256, 0, m->stream>>>(a);
  if (S > 1) k_splitk<EPI><<<blocks((size_t)a.M * a.N, 256), 256, 0, m->stream>>>(a, S);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(LYMPH_MLP_ERR_HIP, std::string("k_gemm: ") + hipGetErrorString(e));
}

// forward of B rows of H[0]; last layer into Y
int forward(LymphMlp* m, int B) {
  for (int l = 0; l < m->L; ++l) {
    GemmArgs a{};
    a.M = B; a.N = m->w[l + 1]; a.K = m->w[l];
    a.A = m->H[l]; a.lda = m->w[l];
    a.B = m->Wb + m->woff[l]; a.ldb = m->w[l];
    a.ldc = m->w[l + 1];
    a.bias = m->P + m->boff[l];
    if (l + 1 < m->L) {
      a.C32 = m->Z[l];
      a.C16 = m->H[l + 1];
      if (int rc = gemm<true, true, EPI_GELU_FWD>(m, a)) return rc;
    } else {
      a.C32 = m->Y;
      if (int rc = gemm<true, true, EPI_F32>(m, a)) return rc;
    }
  }
  return 0;
}

// loss of Y against T rows idx (device), gradients into G
int backward(LymphMlp* m, int B, const float* T, const int* idx, float* loss_out) {
  const int NL = m->w[m->L];
  const size_t ny = (size_t)B * NL;
  const unsigned nb = blocks(ny, kLossPer);
  k_loss<<<nb, 256, 0, m->stream>>>(m->Y, NL, T, idx, B, 2.0f / (float)ny, m->dY, m->part);
  k_sum<<<1, 256, 0, m->stream>>>(m->part, (int)nb, 1.0f / (float)ny, loss_out);
  bf16_t* dZ = m->dY;
  for (int l = m->L - 1; l >= 0; --l) {
    const int N = m->w[l + 1], K = m->w[l];
    GemmArgs a{};  // dW_l = dZ_l^T H_l
    a.M = N; a.N = K; a.K = B;
    a.A = dZ; a.lda = N;
    a.B = m->H[l]; a.ldb = K;
    a.C32 = m->G + m->woff[l]; a.ldc = K;
    if (int rc = gemm<false, false, EPI_F32>(m, a)) return rc;
    k_colsum<<<blocks(N, 64), 256, 0, m->stream>>>(dZ, B, N, m->G + m->boff[l]);
    if (l > 0) {  // dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1})
      bf16_t* nxt = (dZ == m->dZa) ? m->dZb : m->dZa;
      GemmArgs d{};
      d.M = B; d.N = K; d.K = N;
      d.A = dZ; d.lda = N;
      d.B = m->Wb + m->woff[l]; d.ldb = K;
      d.C16 = nxt; d.ldc = K;
      d.aux = m->Z[l - 1];
      if (int rc = gemm<true, false, EPI_GELU_BWD>(m, d)) return rc;
      dZ = nxt;
    }
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(LYMPH_MLP_ERR_HIP, std::string("backward: ") + hipGetErrorString(e));
}

int need(const LymphMlp* m) { return m ? 0 : fail(LYMPH_MLP_ERR_ARG, "null network"); }

}  // namespace

extern "C" {

const char* lymph_mlp_last_error(void) { return g_err.c_str(); }

int lymph_mlp_destroy(LymphMlp* m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : m->owned) (void)hipFree(p);
  if (m->ws) (void)hipFree(m->ws);
  if (m->dorder) (void)hipFree(m->dorder);
  if (m->dloss) (void)hipFree(m->dloss);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int lymph_mlp_create(int n_layers, const int32_t* widths, int max_batch, LymphMlp** out) {
  if (!out) return fail(LYMPH_MLP_ERR_ARG, "out is null");
  *out = nullptr;
  if (n_layers < 1 || !widths) return fail(LYMPH_MLP_ERR_ARG, "need at least one layer");
  if (max_batch < 8 || max_batch % 8) return fail(LYMPH_MLP_ERR_ARG, "max_batch must be a positive multiple of 8");
  for (int l = 0; l <= n_layers; ++l)
    if (widths[l] < 8 || widths[l] % 8) return fail(LYMPH_MLP_ERR_ARG, "widths must be positive multiples of 8");
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0)
    return fail(LYMPH_MLP_ERR_HIP, "no HIP device: liblymph_mlp needs an MI355X (gfx950)");
  auto* m = new LymphMlp;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return fail(LYMPH_MLP_ERR_HIP, "hipStreamCreate failed");
  }
  m->L = n_layers;
  m->max_batch = max_batch;
  m->w.assign(widths, widths + n_layers + 1);
  size_t off = 0;
  int maxh = 8;
  for (int l = 0; l < n_layers; ++l) {
    m->woff.push_back(off);
    off += (size_t)m->w[l + 1] * m->w[l];
    m->boff.push_back(off);
    off += (size_t)m->w[l + 1];
    if (l + 1 < n_layers) maxh = std::max(maxh, m->w[l + 1]);
  }
  m->n_params = off;
  int rc = 0;
  auto al = [&](auto** p, size_t n) {
    if (!rc) rc = dalloc(m, p, n);
  };
  al(&m->P, off);
  al(&m->G, off);
  al(&m->Mo, off);
  al(&m->Vo, off);
  al(&m->Wb, off);
  m->H.assign(n_layers, nullptr);
  m->Z.assign(n_layers, nullptr);
  for (int l = 0; l < n_layers; ++l) al(&m->H[l], (size_t)max_batch * m->w[l]);
  for (int l = 0; l + 1 < n_layers; ++l) al(&m->Z[l], (size_t)max_batch * m->w[l + 1]);
  const int NL = m->w[n_layers];
  al(&m->Y, (size_t)max_batch * NL);
  al(&m->dY, (size_t)max_batch * NL);
  al(&m->dZa, (size_t)max_batch * maxh);
  al(&m->dZb, (size_t)max_batch * maxh);
  al(&m->part, blocks((size_t)max_batch * NL, 256));
  al(&m->loss, 1);
  al(&m->idx, (size_t)max_batch);
  if (rc) {
    lymph_mlp_destroy(m);
    return rc;
  }
  if (hipMemsetAsync(m->P, 0, off * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->Mo, 0, off * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->Vo, 0, off * 4, m->stream) != hipSuccess ||
      hipMemsetAsync(m->Wb, 0, off * 2, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess) {
    lymph_mlp_destroy(m);
    return fail(LYMPH_MLP_ERR_HIP, "hipMemset failed");
  }
  *out = m;
  return 0;
}

int lymph_mlp_n_params(const LymphMlp* m, int64_t* n) {
  if (int rc = need(m)) return rc;
  if (n) *n = (int64_t)m->n_params;
  return 0;
}

int lymph_mlp_set_params(LymphMlp* m, const float* p) {
  if (int rc = need(m)) return rc;
  if (!p) return fail(LYMPH_MLP_ERR_ARG, "null params");
  HIP_TRY(hipMemcpyAsync(m->P, p, m->n_params * 4, hipMemcpyHostToDevice, m->stream));
  k_to_bf16<<<blocks(m->n_params, 256), 256, 0, m->stream>>>(m->P, m->Wb, m->n_params);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(m->Mo, 0, m->n_params * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->Vo, 0, m->n_params * 4, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  return 0;
}

int lymph_mlp_get_params(const LymphMlp* m, float* p) {
  if (int rc = need(m)) return rc;
  if (!p) return fail(LYMPH_MLP_ERR_ARG, "null params");
  HIP_TRY(hipMemcpyAsync(p, m->P, m->n_params * 4, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

int lymph_mlp_set_data(LymphMlp* m, int n, const float* x, const float* y) {
  if (int rc = need(m)) return rc;
  if (n < 1 || !x || !y) return fail(LYMPH_MLP_ERR_ARG, "empty data set");
  const size_t nx = (size_t)n * m->w[0], ny = (size_t)n * m->w[m->L];
  (void)hipStreamSynchronize(m->stream);
  for (float** q : {&m->X, &m->T})
    if (*q) {
      for (auto& o : m->owned)
        if (o == (void*)*q) o = nullptr;
      (void)hipFree(*q);
      *q = nullptr;
    }
  if (int rc = dalloc(m, &m->X, nx)) return rc;
  if (int rc = dalloc(m, &m->T, ny)) return rc;
  HIP_TRY(hipMemcpyAsync(m->X, x, nx * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->T, y, ny * 4, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->n_data = n;
  return 0;
}

int lymph_mlp_train(LymphMlp* m, int n_steps, int batch, const int32_t* order, float lr,
                    float beta1, float beta2, float eps, float* losses) {
  if (int rc = need(m)) return rc;
  if (!m->X) return fail(LYMPH_MLP_ERR_STATE, "no training data: call lymph_mlp_set_data");
  if (batch < 8 || batch > m->max_batch || batch % 8)
    return fail(LYMPH_MLP_ERR_ARG, "batch must be a multiple of 8 and at most max_batch");
  if (n_steps < 0 || !order) return fail(LYMPH_MLP_ERR_ARG, "bad steps/order");
  for (size_t i = 0; i < (size_t)n_steps * batch; ++i)
    if (order[i] < 0 || order[i] >= m->n_data) return fail(LYMPH_MLP_ERR_ARG, "sample index out of range");
  const size_t n_order = std::max<size_t>(1, (size_t)n_steps * batch);
  const size_t n_loss = std::max<size_t>(1, (size_t)n_steps);
  if (n_order > m->dorder_cap || n_loss > m->dloss_cap) {  // grown, then reused by later calls
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (m->dorder) (void)hipFree(m->dorder);
    if (m->dloss) (void)hipFree(m->dloss);
    m->dorder = nullptr;
    m->dloss = nullptr;
    m->dorder_cap = m->dloss_cap = 0;
    HIP_TRY(hipMalloc(&m->dorder, sizeof(int) * n_order));
    HIP_TRY(hipMalloc(&m->dloss, sizeof(float) * n_loss));
    m->dorder_cap = n_order;
    m->dloss_cap = n_loss;
  }
  int* dorder = m->dorder;
  float* dloss = m->dloss;
  HIP_TRY(hipMemcpyAsync(dorder, order, sizeof(int) * (size_t)n_steps * batch, hipMemcpyHostToDevice,
                         m->stream));
  int rc = 0;
  for (int s = 0; s < n_steps && !rc; ++s) {
    const int* idx = dorder + (size_t)s * batch;
    k_gather<<<blocks((size_t)batch * m->w[0], 256), 256, 0, m->stream>>>(m->X, m->w[0], idx, batch, m->H[0]);
    rc = forward(m, batch);
    if (!rc) rc = backward(m, batch, m->T, idx, dloss + s);
    if (!rc) {
      m->adam_t++;
      const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)m->adam_t));
      const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)m->adam_t));
      k_adam<<<blocks(m->n_params, 256), 256, 0, m->stream>>>(m->P, m->Mo, m->Vo, m->G, m->Wb, m->n_params, lr,
                                                beta1, beta2, eps, c1, c2);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail(LYMPH_MLP_ERR_HIP, std::string("k_adam: ") + hipGetErrorString(e));
    }
  }
  if (!rc && losses && n_steps > 0 &&
      hipMemcpyAsync(losses, dloss, sizeof(float) * n_steps, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "loss copy failed");
  if (hipStreamSynchronize(m->stream) != hipSuccess && !rc) rc = fail(LYMPH_MLP_ERR_HIP, "training failed");
  return rc;
}

int lymph_mlp_gradients(LymphMlp* m, int batch, const float* x, const float* y, float* grads,
                        float* loss) {
  if (int rc = need(m)) return rc;
  if (batch < 8 || batch > m->max_batch || batch % 8)
    return fail(LYMPH_MLP_ERR_ARG, "batch must be a multiple of 8 and at most max_batch");
  if (!x || !y) return fail(LYMPH_MLP_ERR_ARG, "null batch");
  const size_t nx = (size_t)batch * m->w[0], ny = (size_t)batch * m->w[m->L];
  float *dx = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc(&dx, nx * 4));
  if (hipMalloc(&dy, ny * 4) != hipSuccess) {
    (void)hipFree(dx);
    return fail(LYMPH_MLP_ERR_HIP, "hipMalloc failed");
  }
  std::vector<int> id(batch);
  for (int i = 0; i < batch; ++i) id[i] = i;
  int rc = 0;
  if (hipMemcpyAsync(dx, x, nx * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(dy, y, ny * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
      hipMemcpyAsync(m->idx, id.data(), 4 * batch, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "upload failed");
  if (!rc) {
    k_gather<<<blocks(nx, 256), 256, 0, m->stream>>>(dx, m->w[0], m->idx, batch, m->H[0]);
    rc = forward(m, batch);
  }
  if (!rc) rc = backward(m, batch, dy, m->idx, m->loss);
  if (!rc && grads &&
      hipMemcpyAsync(grads, m->G, m->n_params * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "gradient copy failed");
  if (!rc && loss && hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "loss copy failed");
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

int lymph_mlp_predict(LymphMlp* m, int n, const float* x, float* y) {
  if (int rc = need(m)) return rc;
  if (n < 0 || (n > 0 && (!x || !y))) return fail(LYMPH_MLP_ERR_ARG, "bad inputs");
  const int w0 = m->w[0], NL = m->w[m->L];
  float* dx = nullptr;
  HIP_TRY(hipMalloc(&dx, (size_t)std::max(1, n) * w0 * 4));
  int rc = 0;
  if (n > 0 &&
      hipMemcpyAsync(dx, x, (size_t)n * w0 * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "upload failed");
  std::vector<int> id(m->max_batch);
  for (int i = 0; i < m->max_batch; ++i) id[i] = i;
  if (!rc &&
      hipMemcpyAsync(m->idx, id.data(), 4 * m->max_batch, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = fail(LYMPH_MLP_ERR_HIP, "upload failed");
  for (int b0 = 0; b0 < n && !rc; b0 += m->max_batch) {
    const int B = std::min(m->max_batch, n - b0);
    k_gather<<<blocks((size_t)B * w0, 256), 256, 0, m->stream>>>(dx + (size_t)b0 * w0, w0, m->idx, B, m->H[0]);
    rc = forward(m, B);
    if (!rc && hipMemcpyAsync(y + (size_t)b0 * NL, m->Y, (size_t)B * NL * 4, hipMemcpyDeviceToHost,
                              m->stream) != hipSuccess)
      rc = fail(LYMPH_MLP_ERR_HIP, "download failed");
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(dx);
  return rc;
}

int lymph_mlp_time_predict(LymphMlp* m, int n, int reps, double* ms) {
  if (int rc = need(m)) return rc;
  if (!m->X || n < 1 || n > m->n_data || reps < 1) return fail(LYMPH_MLP_ERR_ARG, "bad n/reps or no data");
  std::vector<int> id(m->max_batch);
  for (int i = 0; i < m->max_batch; ++i) id[i] = i;
  HIP_TRY(hipMemcpyAsync(m->idx, id.data(), 4 * m->max_batch, hipMemcpyHostToDevice, m->stream));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = 0;
  HIP_TRY(hipEventRecord(e0, m->stream));
  for (int r = 0; r < reps && !rc; ++r)
    for (int b0 = 0; b0 < n && !rc; b0 += m->max_batch) {
      const int B = std::min(m->max_batch, n - b0);
      k_gather<<<blocks((size_t)B * m->w[0], 256), 256, 0, m->stream>>>(m->X + (size_t)b0 * m->w[0], m->w[0],
                                                          m->idx, B, m->H[0]);
      rc = forward(m, B);
    }
  HIP_TRY(hipEventRecord(e1, m->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = f / reps;
  return rc;
}

}  // extern "C"
