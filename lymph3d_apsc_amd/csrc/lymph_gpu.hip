// lymph_gpu.hip -- context management, host-side block pattern and the C ABI
// declared in include/lymph_gpu.h.  All numerics run in the kernels of
// lymph_kernels.hpp; the host only builds index arrays and quadrature tables
// and evaluates the scalar time factor g(t).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lymph_gpu.h"
#include "lymph_kernels.hpp"

using namespace lymph;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return fail(LYMPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));      \
  } while (0)

// HIP_TRY that runs a clean-up first: `handler(code)` returns the code
#define HIP_TRY_OR(expr, handler)                                                         \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return handler(fail(LYMPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e))); \
  } while (0)

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(LYMPH_ERR_HIP, "hipMalloc(" + std::to_string(n * sizeof(T)) + " B): " +
                                   hipGetErrorString(e));
  return 0;
}

// host copy of a device array (set-up paths only)
template <class T>
int d2h(std::vector<T>& h, const T* d, size_t n) {
  h.resize(n);
  if (n) HIP_TRY(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return 0;
}

template <class T>
int upload(T** p, const T* h, size_t n) {
  int rc = dalloc(p, n);
  if (rc) return rc;
  if (n) HIP_TRY(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Staging of lymph_add_point_sources_noalloc: one file-scope device array,
// so that the entry point allocates nothing.  Every GPU the library runs on
// holds its own copy until the process ends; host code never passes the
// array's name (a host-side stand-in) to a kernel or a copy, only the device
// address hipGetSymbolAddress returned for the current device.  Layout:
// xyz[3n] doubles, force[3n] doubles, then tet[n] int32.
constexpr int kNoAllocMaxSources = 16384;
constexpr size_t kScratchBytes = (size_t)kNoAllocMaxSources * (6 * sizeof(double) + sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on device `dev` (which must be current),
// resolved once per device.
int scratch_ptr(int dev, unsigned char** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return fail(LYMPH_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LYMPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return 0;
}

}  // namespace

struct LymphCtx {
  int p = 0, Np = 0, n3 = 0, n3c = 0, nq3 = 0, nq2 = 0;
  double alpha = 10.0, theta = -1.0;
  int n_tet = 0, n_poly = 0, n_blk = 0;
  int layout = LYMPH_LAYOUT_POLY;
  int step_variant = 0;
  int apply_bits = 64;
  bool keep_adg = true;
  int build_mode = LYMPH_BUILD_REFERENCE;
  double ricker_f0 = 5.0, ricker_t0 = 1.2 / 5.0;
  // per-group Ricker wavelets (ensembles): group of each polyhedron, f0/t0
  // per group, and the table of this step's factors
  int n_groups = 0;
  int* gsel = nullptr;
  double *gf0 = nullptr, *gt0 = nullptr, *gtab = nullptr;
  float* K32 = nullptr;     // 32-bit apply: ~K = R^-T K R^-1 (mass-orthonormal basis)
  _Float16* K16 = nullptr;  // 48-bit apply: per entry of K32 its 16-bit correction
  double* tv[5] = {};       // 32-bit stepping: ~u rotation (0..2), ~F_s (3), ~F_d (4)
  double* Ks = nullptr;     // face pairs stored once (lymph_set_symmetric_storage)
  int* sidx = nullptr;      // per block of the full rows: stored block, ~ = transposed
  int* tdst = nullptr;      // stored pair block -> its (Q, P) slot of Q's row, else -1
  int* ssrc = nullptr;      // full-row slot of each stored block (gather list)
  int4* pinfo = nullptr;    // per stored pair block: full-row slot, P, Q, slot of (Q, P)
  double* Sinv = nullptr;   // R^-1 of each mass block, packed like R (pairs-once FP64 update)
  float* Ks32 = nullptr;    // the stored blocks of the 32-bit operator (p = 3)
  _Float16* Ks16 = nullptr; // and their 48-bit corrections, once K16 exists
  size_t ks32_bytes = 0;    // Ks32 and Ks16
  bool sym32_wanted = false;  // LYMPH_SYM_FP64_FP32: the 32- and 48-bit applies store pairs once too
  double* prod = nullptr;   // per-block products of one step (n_blk x 3Np)
  double* rowp = nullptr;   // 32/48-bit pairs once: per polyhedron, its stored blocks' row products summed
  size_t ks_bytes = 0;
  int* blk_row = nullptr;   // row polyhedron of each block
  int* order = nullptr;     // step dispatch order (step_grid(n_poly) entries)
  int max_row = 0;          // most blocks in one block row
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_records = 0;
  double* graph_slot[LYMPH_NUM_SLOTS] = {};  // slot pointers as the captured solve leaves them
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;  // the HIP device current at lymph_ctx_create
  bool built = false;
  std::vector<int> h_row_ptr, h_col, h_tet_poly;
  // device
  MeshDev md{};
  std::vector<void*> owned;  // everything hipMalloc'ed for the mesh
  double *K = nullptr, *A = nullptr, *M = nullptr, *Mm = nullptr, *RM = nullptr, *RMm = nullptr;
  double* slot[LYMPH_NUM_SLOTS] = {};
  double* stage[4] = {};
  double* part = nullptr;
  double* dsum = nullptr;
  int* status = nullptr;
  size_t bytes = 0;
  // receivers
  int n_recv = 0, max_rec = 0, n_rec = 0;
  int* rpoly = nullptr;
  double* rxyz = nullptr;
  double* rnow = nullptr;  // 3 n_recv values of one lymph_sample_receivers call
  int* rpoly_u = nullptr;  // distinct receiver polyhedra
  int n_rpoly_u = 0;
  double* trace = nullptr;
};

namespace {

size_t vec_n(const LymphCtx* c) { return (size_t)c->n_poly * c->n3; }
size_t blk_sz(const LymphCtx* c) { return (size_t)c->n3 * c->n3c; }

template <class F>
int dispatch_p(int p, F&& f) {
  switch (p) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
  }
  return fail(LYMPH_ERR_ARG, "polynomial degree must be 1..4");
}

// Step-kernel variants per degree: (column pairs per unrolled load group,
// neighbour-coefficient broadcast mode); index 0 is the default until
// lymph_autotune picks one.
// Step-kernel variants per degree: column pairs per unrolled load group,
// neighbour-coefficient broadcast mode (0 scalar loads, 1 lane broadcast,
// 2 packed: 64/3Np polyhedra per wave, coefficients through LDS), waves per
// polyhedron; index 0 is the default until lymph_autotune picks one.
template <int P> struct StepVar;
template <> struct StepVar<1> {
  static constexpr int unr[] = {6, 6, 3, 2, 6, 1, 1, 1, 1};
  static constexpr int xs[] = {1, 0, 1, 1, 1, 2, 3, 2, 3};
  static constexpr int wpp[] = {1, 1, 1, 1, 2, 1, 1, 4, 4};
};
template <> struct StepVar<2> {
  static constexpr int unr[] = {5, 5, 15, 3, 15, 5, 5, 1, 1, 1, 1};
  static constexpr int xs[] = {1, 0, 1, 1, 0, 1, 1, 2, 3, 2, 3};
  static constexpr int wpp[] = {1, 1, 1, 1, 1, 2, 4, 1, 1, 4, 4};
};
template <> struct StepVar<3> {
  static constexpr int unr[] = {10, 10, 5, 6, 15, 30, 30, 3, 30, 10, 30, 15};
  static constexpr int xs[] = {1, 0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1};
  static constexpr int wpp[] = {1, 1, 1, 1, 1, 1, 1, 1, 2, 4, 4, 2};
};
template <> struct StepVar<4> {
  static constexpr int unr[] = {4, 4, 2, 8, 1, 4, 4, 4, 2, 8, 8};
  static constexpr int xs[] = {1, 0, 1, 1, 1, 0, 1, 0, 1, 1, 0};
  static constexpr int wpp[] = {1, 1, 1, 1, 1, 2, 2, 4, 4, 4, 4};
};
template <int P> constexpr int n_var() { return (int)(sizeof(StepVar<P>::unr) / sizeof(int)); }

// dynamic LDS of a multi-wave variant (per-block products of the longest
// block row) must stay within this; otherwise variant 0 runs instead
constexpr size_t kMaxDynLds = 64 * 1024;

template <int P, int MODE, int I = 0>
void launch_leapfrog(int variant, int grid, hipStream_t st, const double* K, const int* rp,
                     const int* col, const double* Rp, const double* rs, const double* rd,
                     const double* u, const double* wa, double* out, double g, double c1,
                     double c2, const int* gsel, const double* gtab, const int* ord, size_t dyn) {
  if constexpr (I < n_var<P>()) {
    if (variant == I) {
      constexpr int W = StepVar<P>::wpp[I];
      if constexpr (StepVar<P>::xs[I] == 2) {
        constexpr int G = 64 / (3 * np_of(P));
        const int groups = (grid + G - 1) / G;
        k_leapfrog_packed<P, MODE, 64, W><<<step_grid((groups + W - 1) / W), 64 * W, 0, st>>>(
            K, nullptr, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid);
        return;
      } else if constexpr (StepVar<P>::xs[I] == 3) {
        constexpr int G = 64 / (3 * np_of(P));
        if ((size_t)W * G * dyn > kMaxDynLds) {
          launch_leapfrog<P, MODE, 0>(0, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2,
                                      gsel, gtab, ord, dyn);
          return;
        }
        const int groups = (grid + G - 1) / G;
        k_leapfrog_joint<P, MODE, 64, W><<<step_grid((groups + W - 1) / W), 64 * W, W * G * dyn, st>>>(
            K, nullptr, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid, (int)(G * dyn / 8));
        return;
      } else {
      if (W > 1 && dyn > kMaxDynLds) {
        launch_leapfrog<P, MODE, 0>(0, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2,
                                    gsel, gtab, ord, dyn);
        return;
      }
      k_leapfrog<P, MODE, StepVar<P>::unr[I], StepVar<P>::xs[I], W>
          <<<step_grid(grid), 64 * W, W > 1 ? dyn : 0, st>>>(K, rp, col, Rp, rs, rd, u, wa, out, g,
                                                             c1, c2, gsel, gtab, grid, ord);
      return;
      }
    }
    launch_leapfrog<P, MODE, I + 1>(variant, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2,
                                    gsel, gtab, ord, dyn);
  }
}

template <int P> struct StepVar32;
template <> struct StepVar32<1> {
  static constexpr int unr[] = {3, 1, 3, 1, 1, 1, 1};
  static constexpr int wpp[] = {1, 1, 2, 1, 1, 4, 4};
  static constexpr int pk[] = {0, 0, 0, 1, 2, 1, 2};
};
template <> struct StepVar32<2> {
  static constexpr int unr[] = {4, 2, 8, 4, 4, 1, 1, 1, 1};
  static constexpr int wpp[] = {1, 1, 1, 2, 4, 1, 1, 4, 4};
  static constexpr int pk[] = {0, 0, 0, 0, 0, 1, 2, 1, 2};
};
template <> struct StepVar32<3> {
  static constexpr int unr[] = {5, 3, 15, 5, 5, 15, 15, 8};
  static constexpr int wpp[] = {1, 1, 1, 2, 4, 2, 4, 4};
  static constexpr int pk[] = {0, 0, 0, 0, 0, 0, 0, 0};
};
template <> struct StepVar32<4> {
  static constexpr int unr[] = {3, 1, 9, 3, 3, 9, 9, 13};
  static constexpr int wpp[] = {1, 1, 1, 2, 4, 2, 4, 4};
  static constexpr int pk[] = {0, 0, 0, 0, 0, 0, 0, 0};
};
template <int P> constexpr int n_var32() { return (int)(sizeof(StepVar32<P>::unr) / sizeof(int)); }
// 48-bit apply (k_leapfrog48): column quads per load group, waves per
// polyhedron (pk = 1: the packed step, 64/3Np polyhedra per wave, waves per
// workgroup; pk = 2: the same with the joint block walk; p <= 2)
template <int P> struct StepVar48;
template <> struct StepVar48<1> {  // quads per load group: even (the corrections come in quad pairs)
  static constexpr int unr[] = {4, 4, 4, 1, 1, 1, 1};
  static constexpr int wpp[] = {1, 2, 4, 1, 4, 1, 4};
  static constexpr int pk[] = {0, 0, 0, 1, 1, 2, 2};
};
template <> struct StepVar48<2> {
  static constexpr int unr[] = {8, 8, 8, 1, 1, 1, 1};
  static constexpr int wpp[] = {1, 2, 4, 1, 4, 1, 4};
  static constexpr int pk[] = {0, 0, 0, 1, 1, 2, 2};
};
template <> struct StepVar48<3> {
  static constexpr int unr[] = {6, 16, 6, 16, 6, 16};
  static constexpr int wpp[] = {1, 1, 2, 2, 4, 4};
  static constexpr int pk[] = {0, 0, 0, 0, 0, 0};
};
template <> struct StepVar48<4> {
  static constexpr int unr[] = {14, 14, 14};
  static constexpr int wpp[] = {1, 2, 4};
  static constexpr int pk[] = {0, 0, 0};
};
template <int P> constexpr int n_var48() { return (int)(sizeof(StepVar48<P>::unr) / sizeof(int)); }
// lymph_autotune reports into caller arrays of 16 entries
static_assert(n_var<1>() <= 16 && n_var<2>() <= 16 && n_var<3>() <= 16 && n_var<4>() <= 16 &&
                  n_var32<1>() <= 16 && n_var32<2>() <= 16 && n_var32<3>() <= 16 &&
                  n_var32<4>() <= 16 && n_var48<1>() <= 16 && n_var48<2>() <= 16 &&
                  n_var48<3>() <= 16 && n_var48<4>() <= 16,
              "at most 16 step variants per degree");
template <int P> int n_var_of(int bits) {
  return bits == 64 ? n_var<P>() : (bits == 32 ? n_var32<P>() : n_var48<P>());
}

template <int P, int MODE, int I = 0>
void launch_leapfrog32(int variant, int grid, hipStream_t st, const float* K, const int* rp,
                       const int* col, const double* Rp, const double* rs, const double* rd,
                       const double* u, const double* wa, double* out, double g, double c1,
                       double c2, const int* gsel, const double* gtab, const int* ord, size_t dyn) {
  if constexpr (I < n_var32<P>()) {
    if (variant == I) {
      constexpr int W = StepVar32<P>::wpp[I];
      if constexpr (StepVar32<P>::pk[I] == 1) {
        constexpr int G = 64 / (3 * np_of(P));
        const int groups = (grid + G - 1) / G;
        k_leapfrog_packed<P, MODE, 32, W><<<step_grid((groups + W - 1) / W), 64 * W, 0, st>>>(
            K, nullptr, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid);
        return;
      } else if constexpr (StepVar32<P>::pk[I] == 2) {
        constexpr int G = 64 / (3 * np_of(P));
        if ((size_t)W * G * dyn > kMaxDynLds) {
          launch_leapfrog32<P, MODE, 0>(0, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2,
                                        gsel, gtab, ord, dyn);
          return;
        }
        const int groups = (grid + G - 1) / G;
        k_leapfrog_joint<P, MODE, 32, W><<<step_grid((groups + W - 1) / W), 64 * W, W * G * dyn, st>>>(
            K, nullptr, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid, (int)(G * dyn / 8));
        return;
      } else {
      if (W > 1 && dyn > kMaxDynLds) {
        launch_leapfrog32<P, MODE, 0>(0, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1, c2,
                                      gsel, gtab, ord, dyn);
        return;
      }
      k_leapfrog32<P, MODE, StepVar32<P>::unr[I], W>
          <<<step_grid(grid), 64 * W, W > 1 ? dyn : 0, st>>>(K, rp, col, Rp, rs, rd, u, wa, out, g,
                                                             c1, c2, gsel, gtab, grid, ord);
      return;
      }
    }
    launch_leapfrog32<P, MODE, I + 1>(variant, grid, st, K, rp, col, Rp, rs, rd, u, wa, out, g, c1,
                                      c2, gsel, gtab, ord, dyn);
  }
}

template <int P, int MODE, int I = 0>
void launch_leapfrog48(int variant, int grid, hipStream_t st, const float* K, const _Float16* Kr,
                       const int* rp, const int* col, const double* rs, const double* rd,
                       const double* u, const double* wa, double* out, double g, double c1,
                       double c2, const int* gsel, const double* gtab, const int* ord, size_t dyn) {
  if constexpr (I < n_var48<P>()) {
    if (variant == I) {
      constexpr int W = StepVar48<P>::wpp[I];
      if constexpr (StepVar48<P>::pk[I] == 1) {
        static_assert(3 * np_of(P) <= 32, "packed step needs 3Np <= 32");
        constexpr int G = 64 / (3 * np_of(P));
        const int groups = (grid + G - 1) / G;
        k_leapfrog_packed<P, MODE, 48, W><<<step_grid((groups + W - 1) / W), 64 * W, 0, st>>>(
            K, Kr, rp, col, nullptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid);
        return;
      } else if constexpr (StepVar48<P>::pk[I] == 2) {
        static_assert(3 * np_of(P) <= 32, "joint packed step needs 3Np <= 32");
        constexpr int G = 64 / (3 * np_of(P));
        if ((size_t)W * G * dyn > kMaxDynLds) {
          launch_leapfrog48<P, MODE, 0>(0, grid, st, K, Kr, rp, col, rs, rd, u, wa, out, g, c1, c2,
                                        gsel, gtab, ord, dyn);
          return;
        }
        const int groups = (grid + G - 1) / G;
        k_leapfrog_joint<P, MODE, 48, W><<<step_grid((groups + W - 1) / W), 64 * W, W * G * dyn, st>>>(
            K, Kr, rp, col, nullptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid, (int)(G * dyn / 8));
        return;
      } else {
        if (W > 1 && dyn > kMaxDynLds) {
          launch_leapfrog48<P, MODE, 0>(0, grid, st, K, Kr, rp, col, rs, rd, u, wa, out, g, c1, c2,
                                        gsel, gtab, ord, dyn);
          return;
        }
        k_leapfrog48<P, MODE, StepVar48<P>::unr[I], W><<<step_grid(grid), 64 * W, W > 1 ? dyn : 0, st>>>(
            K, Kr, rp, col, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, grid, ord);
        return;
      }
    }
    launch_leapfrog48<P, MODE, I + 1>(variant, grid, st, K, Kr, rp, col, rs, rd, u, wa, out, g, c1,
                                      c2, gsel, gtab, ord, dyn);
  }
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LYMPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

unsigned grid1(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
size_t Dims16Any(int p) {  // Dims16<p>::BS
  const size_t n3 = (size_t)(p + 1) * (p + 2) * (p + 3) / 2;
  const size_t nq = n3 / 4;
  return ((nq / 2) * n3 * 8 + (nq % 2) * n3 * 4 + n3 * (n3 % 4) + 7) / 8 * 8;
}
size_t Dims32Any(int p) {  // Dims32<p>::BS
  const size_t n3 = (size_t)(p + 1) * (p + 2) * (p + 3) / 2;
  return (n3 * n3 + 3) / 4 * 4;
}


double g_of(const LymphCtx* c, int problem, double t) {
  if (problem == LYMPH_PROB_RICKER) return ricker(c->ricker_f0, c->ricker_t0, t);
  return time_factor(problem, t);
}

// host vector (user layout) -> device staging (poly layout)
int h2d_vec(LymphCtx* c, const double* h, double* d) {
  if (!h) return fail(LYMPH_ERR_ARG, "null vector");
  const size_t n = vec_n(c);
  if (c->layout == LYMPH_LAYOUT_POLY) {
    HIP_TRY(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(c->stage[3], h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    k_to_poly<<<grid1(n, 256), 256, 0, c->stream>>>(c->stage[3], d, c->n_poly, c->n3);
    if (int rc = check_launch("k_to_poly")) return rc;
  }
  return 0;
}

int d2h_vec(LymphCtx* c, const double* d, double* h) {
  if (!h) return fail(LYMPH_ERR_ARG, "null vector");
  const size_t n = vec_n(c);
  if (c->layout == LYMPH_LAYOUT_POLY) {
    HIP_TRY(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    k_to_fortran<<<grid1(n, 256), 256, 0, c->stream>>>(d, c->stage[3], c->n_poly, c->n3);
    if (int rc = check_launch("k_to_fortran")) return rc;
    HIP_TRY(hipMemcpyAsync(h, c->stage[3], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int launch_step(LymphCtx* c, int mode, double g, double c1, double c2, const double* u,
                const double* wa, double* out, bool grouped = false) {
  const int* gsel = grouped ? c->gsel : nullptr;
  const double* gtab = grouped ? c->gtab : nullptr;
  const size_t dyn = (size_t)c->max_row * c->n3 * sizeof(double);
  return dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    // the 32-bit apply works on mass-orthonormal vectors: ~F from tilde_rhs
    const double* rs = c->apply_bits != 64 ? c->tv[3] : c->slot[LYMPH_SLOT_RHS_STAT];
    const double* rd = c->apply_bits != 64 ? c->tv[4] : c->slot[LYMPH_SLOT_RHS_DYN];
    if (c->apply_bits == 32 && c->Ks32) {
      if constexpr (3 * np_of(P) % 12 == 0 && 3 * np_of(P) <= 64) {
        float* prod32 = reinterpret_cast<float*>(c->prod);  // the 32-bit products in the same buffer
        k_sym_products32<P><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
            c->Ks32, c->md.row_ptr, c->md.col, c->sidx, c->tdst, u, prod32, c->rowp, c->n_poly, c->order);
        if (int rc = check_launch("k_sym_products32")) return rc;
        if (mode == 0)
          k_sym_update32<P, 0><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              prod32, c->rowp, c->sidx, c->md.row_ptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
        else
          k_sym_update32<P, 1><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              prod32, c->rowp, c->sidx, c->md.row_ptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
      } else {
        return fail(LYMPH_ERR_STATE, "32-bit stored pairs need 3Np a multiple of 12");
      }
    } else if (c->apply_bits == 48 && c->Ks16) {
      if constexpr (3 * np_of(P) % 12 == 0 && 3 * np_of(P) <= 64) {
        k_sym_products48<P><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
            c->Ks32, c->Ks16, c->md.row_ptr, c->md.col, c->sidx, c->tdst, u, c->prod, c->rowp, c->n_poly,
            c->order);
        if (int rc = check_launch("k_sym_products48")) return rc;
        if (mode == 0)
          k_sym_update32<P, 0, double><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              c->prod, c->rowp, c->sidx, c->md.row_ptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
        else
          k_sym_update32<P, 1, double><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              c->prod, c->rowp, c->sidx, c->md.row_ptr, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
      } else {
        return fail(LYMPH_ERR_STATE, "48-bit stored pairs need 3Np a multiple of 12");
      }
    } else if (c->apply_bits == 48) {
      if (mode == 0)
        launch_leapfrog48<P, 0>(c->step_variant, c->n_poly, c->stream, c->K32, c->K16, c->md.row_ptr,
                                c->md.col, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
      else
        launch_leapfrog48<P, 1>(c->step_variant, c->n_poly, c->stream, c->K32, c->K16, c->md.row_ptr,
                                c->md.col, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
    } else if (c->apply_bits == 32) {
      if (mode == 0)
        launch_leapfrog32<P, 0>(c->step_variant, c->n_poly, c->stream, c->K32, c->md.row_ptr,
                                c->md.col, c->RM, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
      else
        launch_leapfrog32<P, 1>(c->step_variant, c->n_poly, c->stream, c->K32, c->md.row_ptr,
                                c->md.col, c->RM, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
    } else if (c->Ks) {
      if constexpr (3 * np_of(P) <= 64 && (3 * np_of(P)) % 2 == 0) {
        const int n_pair = (c->n_blk - c->n_poly) / 2;
        if constexpr (3 * np_of(P) <= 32) {  // several stored blocks per wave
          constexpr int G = 64 / (3 * np_of(P));
          const int waves = (c->n_poly + n_pair + G - 1) / G;
          k_sym_products_packed<P><<<step_grid(waves), 64, 0, c->stream>>>(
              c->Ks, c->ssrc, c->pinfo, u, c->prod, c->n_poly, c->n_poly + n_pair);
          if (int rc = check_launch("k_sym_products_packed")) return rc;
          const unsigned ug = step_grid((c->n_poly + G - 1) / G);
          if (mode == 0)
            k_sym_update_packed<P, 0><<<ug, 64, 0, c->stream>>>(c->prod, c->md.row_ptr, c->Sinv, rs, rd, u,
                                                                wa, out, g, c1, c2, gsel, gtab, c->n_poly);
          else
            k_sym_update_packed<P, 1><<<ug, 64, 0, c->stream>>>(c->prod, c->md.row_ptr, c->Sinv, rs, rd, u,
                                                                wa, out, g, c1, c2, gsel, gtab, c->n_poly);
          return check_launch("k_sym_update_packed");
        } else {
          const int self_per = (c->n_poly + 7) / 8, pair_per = (n_pair + 7) / 8;
          k_sym_products<P><<<8u * (unsigned)(self_per + pair_per), 64, 0, c->stream>>>(
              c->Ks, c->ssrc, c->pinfo, u, c->prod, c->n_poly, n_pair, self_per, pair_per);
          if (int rc = check_launch("k_sym_products")) return rc;
        }
        if (mode == 0)
          k_sym_update<P, 0><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              c->prod, c->md.row_ptr, c->Sinv, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
        else
          k_sym_update<P, 1><<<step_grid(c->n_poly), 64, 0, c->stream>>>(
              c->prod, c->md.row_ptr, c->Sinv, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->n_poly,
              c->order);
      } else {
        return fail(LYMPH_ERR_STATE, "symmetric storage needs p <= 3");
      }
    } else if (mode == 0) {
      launch_leapfrog<P, 0>(c->step_variant, c->n_poly, c->stream, c->K, c->md.row_ptr, c->md.col,
                            c->RM, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
    } else {
      launch_leapfrog<P, 1>(c->step_variant, c->n_poly, c->stream, c->K, c->md.row_ptr, c->md.col,
                            c->RM, rs, rd, u, wa, out, g, c1, c2, gsel, gtab, c->order, dyn);
    }
    return check_launch("k_leapfrog");
  });
}

// Mass-orthonormal transforms (32-bit apply): mode 0 ~u = R u, 1 u = R^-1 ~u,
// 2 ~F = R^-T F, one wave per polyhedron; `list` restricts them to n listed
// polyhedra.
int tilde(LymphCtx* c, int mode, const double* in, double* out, const int* list = nullptr,
          int n = -1) {
  if (n < 0) n = c->n_poly;
  if (n == 0) return 0;
  return dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if (mode == 0)
      k_tilde<P, 0><<<(unsigned)n, 64, 0, c->stream>>>(c->RM, in, out, list, n);
    else if (mode == 1)
      k_tilde_back<P><<<(unsigned)n, 64, 0, c->stream>>>(c->RM, in, out, list, n);
    else
      k_tilde<P, 2><<<(unsigned)n, 64, 0, c->stream>>>(c->RM, in, out, list, n);
    return check_launch("k_tilde");
  });
}

// Largest max |K_PQ - K_QP^T| / max |K_PQ| of a pair taken as symmetric: the
// two blocks come from the two sides' own face quadratures, which agree to a
// few ulps where the bilinear form is symmetric.
constexpr double kSymTol = 1e-12;

// A captured solve holds the step's buffers by address: drop it whenever they
// change, so that lymph_replay reports "no captured solve" instead of reading
// freed memory.
void drop_graph(LymphCtx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

// Drop the once-stored face pairs (operator rebuilt, storage switched off).
int sym_off(LymphCtx* c) {
  if (!c->Ks && !c->sidx && !c->tdst && !c->prod && !c->ssrc && !c->Ks32 && !c->Ks16 && !c->Sinv &&
      !c->pinfo && !c->rowp)
    return 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
  drop_graph(c);
  if (c->ssrc) HIP_TRY(hipFree(c->ssrc));
  if (c->Ks32) HIP_TRY(hipFree(c->Ks32));
  if (c->Ks16) HIP_TRY(hipFree(c->Ks16));
  if (c->rowp) HIP_TRY(hipFree(c->rowp));
  c->rowp = nullptr;
  c->ssrc = nullptr;
  c->Ks32 = nullptr;
  c->Ks16 = nullptr;
  c->bytes -= c->ks32_bytes;
  c->ks32_bytes = 0;
  if (c->Ks) HIP_TRY(hipFree(c->Ks));
  if (c->sidx) HIP_TRY(hipFree(c->sidx));
  if (c->tdst) HIP_TRY(hipFree(c->tdst));
  if (c->prod) HIP_TRY(hipFree(c->prod));
  if (c->Sinv) HIP_TRY(hipFree(c->Sinv));
  if (c->pinfo) HIP_TRY(hipFree(c->pinfo));
  c->Sinv = nullptr;
  c->pinfo = nullptr;
  c->Ks = nullptr;
  c->sidx = nullptr;
  c->tdst = nullptr;
  c->prod = nullptr;
  c->bytes -= c->ks_bytes;
  c->ks_bytes = 0;
  return 0;
}

int sym32_build(LymphCtx* c);

// Store every face pair once: check K_QP = K_PQ^T pair by pair and K_PP =
// K_PP^T on the device (largest ratio to *asym), then gather the self blocks
// (upper triangles) and the blocks of neighbours Q > P into Ks, index the full
// rows into it, and form S = R^-1 of each mass block for the update.
int sym_on(LymphCtx* c, double tol, double* asym) {
  const int np = c->n_poly;
  drop_graph(c);  // a captured solve would keep stepping the full rows
  const std::vector<int>& rp = c->h_row_ptr;
  const std::vector<int>& cl = c->h_col;
  std::vector<int> partner(c->n_blk, -1), sid(c->n_blk, -1), src, self_s, pb, pt;
  for (int P = 0; P < np; ++P)
    for (int b = rp[P]; b < rp[P + 1]; ++b) {
      const int Q = cl[b];
      if (Q == P) {
        pb.push_back(b);
        pt.push_back(b);
        continue;
      }
      for (int b2 = rp[Q]; b2 < rp[Q + 1]; ++b2)
        if (cl[b2] == P) {
          partner[b] = b2;
          break;
        }
      if (partner[b] < 0) return fail(LYMPH_ERR_STATE, "neighbour rows do not match");
      if (Q > P) {
        pb.push_back(b);
        pt.push_back(partner[b]);
      }
    }
  // stored order: the self blocks first (stored index = polyhedron), then the
  // pair blocks in row order, so that a stored block's kind follows from its
  // index alone
  for (int P = 0; P < np; ++P) {
    int self = -1;
    for (int b = rp[P]; b < rp[P + 1]; ++b)
      if (cl[b] == P) self = b;
    if (self < 0) return fail(LYMPH_ERR_STATE, "a block row without its self block");
    sid[self] = P;
    self_s.push_back(P);
    src.push_back(self);
  }
  std::vector<int4> pi;  // the pair blocks' (slot, P, Q, slot of (Q, P))
  for (int P = 0; P < np; ++P)
    for (int b = rp[P]; b < rp[P + 1]; ++b)
      if (cl[b] > P) {
        sid[b] = (int)src.size();
        src.push_back(b);
        pi.push_back(make_int4(b, P, cl[b], partner[b]));
      }
  std::vector<int> sx(c->n_blk), td(c->n_blk, -1);
  for (int P = 0; P < np; ++P)
    for (int b = rp[P]; b < rp[P + 1]; ++b) {
      sx[b] = cl[b] >= P ? sid[b] : ~sid[partner[b]];
      if (cl[b] > P) td[b] = partner[b];
    }
  const int n_pairs = (int)pb.size();
  double worst = 0.0;
  if (n_pairs > 0) {
    int *dpb = nullptr, *dpt = nullptr;
    double* dr = nullptr;
    int rc = dalloc(&dpb, n_pairs);
    if (!rc) rc = dalloc(&dpt, n_pairs);
    if (!rc) rc = dalloc(&dr, n_pairs);
    std::vector<double> ratio(n_pairs);
    if (!rc && (hipMemcpy(dpb, pb.data(), n_pairs * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dpt, pt.data(), n_pairs * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(LYMPH_ERR_HIP, "copy of the pair lists failed");
    if (!rc)
      rc = dispatch_p(c->p, [&](auto pc) -> int {
        constexpr int P = decltype(pc)::value;
        if constexpr ((3 * np_of(P)) % 2 == 0) {
          k_pair_asym<P><<<(unsigned)n_pairs, 64, 0, c->stream>>>(c->K, dpb, dpt, n_pairs, dr);
          return check_launch("k_pair_asym");
        } else {
          return fail(LYMPH_ERR_STATE, "symmetric storage needs p <= 3");
        }
      });
    if (!rc && (hipMemcpyAsync(ratio.data(), dr, n_pairs * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
      rc = fail(LYMPH_ERR_HIP, "copy of the pair ratios failed");
    (void)hipFree(dpb);
    (void)hipFree(dpt);
    (void)hipFree(dr);
    if (rc) return rc;
    for (double r : ratio)
      if (!(r <= worst)) worst = r;  // a NaN ratio sticks
  }
  if (asym) *asym = worst;
  if (!(worst <= tol))
    return fail(LYMPH_ERR_ARG, "operator is not symmetric: max |K_PQ - K_QP^T| / max |K_PQ| = " +
                                   std::to_string(worst));
  const int n_s = (int)src.size();
  const size_t bs = blk_sz(c);
  int* dsrc = nullptr;
  if (int rc = dalloc(&c->Ks, (size_t)n_s * bs)) return rc;
  int rc = dalloc(&c->sidx, c->n_blk);
  if (!rc) rc = dalloc(&c->tdst, c->n_blk);
  if (!rc) rc = dalloc(&c->prod, (size_t)c->n_blk * prod_stride(c->n3));
  if (!rc) rc = dalloc(&dsrc, n_s);
  if (!rc && !pi.empty()) rc = dalloc(&c->pinfo, pi.size());
  if (!rc && !pi.empty() &&
      hipMemcpy(c->pinfo, pi.data(), pi.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(LYMPH_ERR_HIP, "copy of the pair list failed");
  if (!rc && (hipMemcpy(c->sidx, sx.data(), sx.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(c->tdst, td.data(), td.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(dsrc, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    rc = fail(LYMPH_ERR_HIP, "copy of the block index failed");
  if (!rc) {
    k_gather_blocks<double><<<4096, 256, 0, c->stream>>>(c->K, c->Ks, dsrc, n_s, bs);
    rc = check_launch("k_gather_blocks");
  }
  // the self blocks' upper triangles over their slots' heads, and S = R^-1
  const int n_self = (int)self_s.size();
  const size_t npk = (size_t)c->Np * (c->Np + 1) / 2;
  int* dself = nullptr;
  if (!rc) rc = dalloc(&dself, n_self);
  if (!rc) rc = dalloc(&c->Sinv, (size_t)np * npk);
  if (!rc && hipMemcpy(dself, self_s.data(), n_self * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(LYMPH_ERR_HIP, "copy of the self-block list failed");
  if (!rc)
    rc = dispatch_p(c->p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      if constexpr ((3 * np_of(P)) % 2 == 0) {
        const size_t ne = (size_t)n_self * self_upper_off(3 * np_of(P) / 2);
        k_pack_self_upper<P><<<grid1(ne, 256), 256, 0, c->stream>>>(c->K, c->Ks, dsrc, dself, n_self);
        if (int r = check_launch("k_pack_self_upper")) return r;
        k_inv_upper<P><<<grid1((size_t)np * c->Np, 64), 64, 0, c->stream>>>(c->RM, c->Sinv, np);
        return check_launch("k_inv_upper");
      } else {
        return fail(LYMPH_ERR_STATE, "symmetric storage needs p <= 3");
      }
    });
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(LYMPH_ERR_HIP, "gather failed");
  if (dself) (void)hipFree(dself);
  c->ssrc = dsrc;  // kept for the 32-bit blocks (freed with the stored pairs)
  if (rc) {
    c->ks_bytes = 0;
    (void)sym_off(c);
    return rc;
  }
  c->ks_bytes = (size_t)n_s * bs * sizeof(double) + (2 * (size_t)c->n_blk + n_s) * sizeof(int) +
                (size_t)c->n_blk * prod_stride(c->n3) * sizeof(double) + (size_t)np * npk * sizeof(double) +
                pi.size() * sizeof(int4);
  c->bytes += c->ks_bytes;
  // the 32-bit operator's copy, if asked for and it exists already; a failure
  // leaves no stored pairs switched on
  if (int rc2 = sym32_build(c)) {
    (void)sym_off(c);
    return rc2;
  }
  return 0;
}

// The stored pairs of the 32-bit operator (and of its 48-bit corrections
// once K16 exists), once both it and the stored list exist and
// LYMPH_SYM_FP64_FP32 asked for them (p = 3: whole quads per chunk);
// otherwise the 32- and 48-bit applies keep their full rows.  On failure no
// partial copy is kept.
int sym32_build(LymphCtx* c) {
  if (!c->sym32_wanted || !c->ssrc || !c->K32 || c->p != 3) return 0;
  const int n_s = c->n_poly + (c->n_blk - c->n_poly) / 2;
  if (!c->rowp) {  // each polyhedron's summed row products
    const size_t n = (size_t)c->n_poly * prod_stride(c->n3);
    if (int rc = dalloc(&c->rowp, n)) return rc;
    c->ks32_bytes += n * sizeof(double);
    c->bytes += n * sizeof(double);
  }
  if (!c->Ks32) {
    const size_t bs = (size_t)Dims32Any(c->p);
    float* ks = nullptr;
    if (int rc = dalloc(&ks, (size_t)n_s * bs)) return rc;
    k_gather_blocks<float><<<4096, 256, 0, c->stream>>>(c->K32, ks, c->ssrc, n_s, bs);
    int rc = check_launch("k_gather_blocks");
    if (!rc) {  // the self blocks (stored first) keep their upper triangles
      const size_t ne = (size_t)c->n_poly * self32_off(Dims32<3>::NQ);
      k_pack_self32<3><<<grid1(ne, 256), 256, 0, c->stream>>>(c->K32, ks, c->ssrc, c->n_poly);
      rc = check_launch("k_pack_self32");
    }
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess)
      rc = fail(LYMPH_ERR_HIP, "gather of the 32-bit stored pairs failed");
    if (rc) {
      (void)hipFree(ks);
      return rc;
    }
    c->Ks32 = ks;
    c->ks32_bytes += (size_t)n_s * bs * sizeof(float);
    c->bytes += (size_t)n_s * bs * sizeof(float);
  }
  if (c->K16 && !c->Ks16) {
    const size_t bs = (size_t)Dims16Any(c->p);
    _Float16* ks = nullptr;
    if (int rc = dalloc(&ks, (size_t)n_s * bs)) return rc;
    k_gather_blocks<_Float16><<<4096, 256, 0, c->stream>>>(c->K16, ks, c->ssrc, n_s, bs);
    int rc = check_launch("k_gather_blocks");
    if (!rc) {
      const size_t ne = (size_t)c->n_poly * self16_off(Dims32<3>::N3, Dims16<3>::NPR);
      k_pack_self16<3><<<grid1(ne, 256), 256, 0, c->stream>>>(c->K16, ks, c->ssrc, c->n_poly);
      rc = check_launch("k_pack_self16");
    }
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess)
      rc = fail(LYMPH_ERR_HIP, "gather of the 48-bit stored corrections failed");
    if (rc) {
      (void)hipFree(ks);
      return rc;
    }
    c->Ks16 = ks;
    c->ks32_bytes += (size_t)n_s * bs * sizeof(_Float16);
    c->bytes += (size_t)n_s * bs * sizeof(_Float16);
  }
  return 0;
}

// ~F_s, ~F_d of the current RHS slots (every 32-bit stepping call starts here)
int tilde_rhs(LymphCtx* c) {
  if (int rc = tilde(c, 2, c->slot[LYMPH_SLOT_RHS_STAT], c->tv[3])) return rc;
  return tilde(c, 2, c->slot[LYMPH_SLOT_RHS_DYN], c->tv[4]);
}

// ~K from the FP64 blocks and the mass factors (K32, the tilde vectors and,
// for the 48-bit apply, K16 allocated by the caller)
int build_ortho(LymphCtx* c) {
  if (c->K16)  // the corrections' unused slots (an odd count of quads) read as zeros
    HIP_TRY(hipMemsetAsync(c->K16, 0, (size_t)c->n_blk * Dims16Any(c->p) * sizeof(_Float16), c->stream));
  double* S = nullptr;
  const size_t npk = (size_t)c->Np * (c->Np + 1) / 2;
  if (int rc = dalloc(&S, (size_t)c->n_poly * npk)) return rc;
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_inv_upper<P><<<grid1((size_t)c->n_poly * c->Np, 64), 64, 0, c->stream>>>(c->RM, S, c->n_poly);
    if (int r = check_launch("k_inv_upper")) return r;
    const size_t chunk = 1u << 20;  // workgroups per launch (one per block)
    for (size_t b0 = 0; b0 < (size_t)c->n_blk; b0 += chunk) {
      const unsigned nb = (unsigned)std::min(chunk, (size_t)c->n_blk - b0);
      k_ortho_operator<P><<<nb, 256, 0, c->stream>>>(c->K, S, c->blk_row, c->md.col, c->K32, c->K16,
                                                     b0);
      if (int r = check_launch("k_ortho_operator")) return r;
    }
    return 0;
  });
  const hipError_t e = hipStreamSynchronize(c->stream);
  (void)hipFree(S);
  if (rc) return rc;
  if (e != hipSuccess) return fail(LYMPH_ERR_HIP, "k_ortho_operator failed");
  return 0;
}

// Time factor of `problem` at t for a step launched next: the scalar g, and
// with per-group Ricker wavelets the device table for this step (filled on
// the context's stream, so each step reads its own t).
int prepare_g(LymphCtx* c, int problem, double t, double* g, bool* grouped) {
  *g = g_of(c, problem, t);
  *grouped = problem == LYMPH_PROB_RICKER && c->n_groups > 0;
  if (*grouped) {
    k_ricker_table<<<grid1(c->n_groups, 64), 64, 0, c->stream>>>(c->gf0, c->gt0, c->n_groups, t,
                                                                c->gtab);
    return check_launch("k_ricker_table");
  }
  return 0;
}

int errors_dev(LymphCtx* c, const double* uh, const double* ue, double* l2sq, double* dgsq) {
  const size_t n = vec_n(c);
  double res[2];
  if (dgsq && !c->A) return fail(LYMPH_ERR_STATE, "A_dg blocks were not kept (lymph_ctx_set_keep_adg)");
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if (l2sq) {
      k_l2_terms<P><<<grid1(3 * (size_t)c->n_poly, 128), 128, 0, c->stream>>>(c->Mm, uh, ue, c->part,
                                                                               c->n_poly);
      k_seq_sum<<<1, 64, 0, c->stream>>>(c->part, 3 * c->n_poly, c->dsum);
    }
    if (dgsq) {
      k_sub<<<grid1(n, 256), 256, 0, c->stream>>>(uh, ue, c->stage[2], n);
      launch_leapfrog<P, 2>(0, c->n_poly, c->stream, c->A, c->md.row_ptr, c->md.col, nullptr,
                            nullptr, nullptr, c->stage[2], nullptr, c->part, 0.0, 0.0, 0.0,
                            nullptr, nullptr, c->order, 0);
      k_seq_sum<<<1, 64, 0, c->stream>>>(c->part, c->n_poly, c->dsum + 1);
    }
    return check_launch("error kernels");
  });
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(res, c->dsum, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (l2sq) *l2sq = res[0];
  if (dgsq) *dgsq = res[1];
  return 0;
}

int need_ctx(const LymphCtx* c) {
  if (!c) return fail(LYMPH_ERR_ARG, "null context");
  return 0;
}
int need_built(const LymphCtx* c) {
  if (int rc = need_ctx(c)) return rc;
  if (!c->built) return fail(LYMPH_ERR_STATE, "operator not built: call lymph_build_operator first");
  return 0;
}

}  // namespace

extern "C" {

const char* lymph_last_error(void) { return g_err.c_str(); }
const char* lymph_version(void) { return "lymph3d_apsc_amd 0.1 (gfx950)"; }

int lymph_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

double lymph_time_function(int problem, double t) { return time_factor(problem, t); }

int lymph_ctx_destroy(LymphCtx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->rpoly) (void)hipFree(c->rpoly);
  if (c->rxyz) (void)hipFree(c->rxyz);
  if (c->rnow) (void)hipFree(c->rnow);
  if (c->trace) (void)hipFree(c->trace);
  if (c->K32) (void)hipFree(c->K32);
  if (c->K16) (void)hipFree(c->K16);
  for (double* q : c->tv)
    if (q) (void)hipFree(q);
  if (c->rpoly_u) (void)hipFree(c->rpoly_u);
  for (void* q : {(void*)c->gsel, (void*)c->gf0, (void*)c->gt0, (void*)c->gtab})
    if (q) (void)hipFree(q);
  if (c->K) (void)hipFree(c->K);
  if (c->A) (void)hipFree(c->A);
  if (c->Ks) (void)hipFree(c->Ks);
  if (c->sidx) (void)hipFree(c->sidx);
  if (c->tdst) (void)hipFree(c->tdst);
  if (c->prod) (void)hipFree(c->prod);
  if (c->ssrc) (void)hipFree(c->ssrc);
  if (c->Ks32) (void)hipFree(c->Ks32);
  if (c->Ks16) (void)hipFree(c->Ks16);
  if (c->rowp) (void)hipFree(c->rowp);
  if (c->Sinv) (void)hipFree(c->Sinv);
  if (c->pinfo) (void)hipFree(c->pinfo);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (void* p : c->owned) (void)hipFree(p);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int lymph_ctx_create(const LymphMesh* mesh, int p, double alpha, double theta, LymphCtx** out) {
  if (!out) return fail(LYMPH_ERR_ARG, "out is null");
  *out = nullptr;
  if (!mesh || !mesh->xyz || !mesh->tet_vert || !mesh->neigh || !mesh->normal || !mesh->area ||
      !mesh->bbox || !mesh->hk || !mesh->rho || !mesh->lam || !mesh->mu || !mesh->bc_space)
    return fail(LYMPH_ERR_ARG, "mesh has a null array");
  if (p < 1 || p > kMaxP) return fail(LYMPH_ERR_ARG, "polynomial degree must be 1..4");
  if (mesh->n_tet <= 0 || mesh->n_poly <= 0 || mesh->n_node <= 0)
    return fail(LYMPH_ERR_ARG, "empty mesh");
  const int nt = mesh->n_tet, np = mesh->n_poly;
  std::vector<int> tpoly(nt);
  for (int t = 0; t < nt; ++t) {
    tpoly[t] = mesh->tet_poly ? mesh->tet_poly[t] : t;
    if (tpoly[t] < 0 || tpoly[t] >= np) return fail(LYMPH_ERR_ARG, "tet_poly out of range");
  }
  for (int i = 0; i < 4 * nt; ++i) {
    const int e = mesh->neigh[i];
    if (e >= nt || (e < 0 && e != -1 && e != -2))
      return fail(LYMPH_EXIT_NEIGHBOUR, "neigh entry out of range at " + std::to_string(i));
  }
  for (int i = 0; i < 4 * nt; ++i)
    if (mesh->tet_vert[i] < 0 || mesh->tet_vert[i] >= mesh->n_node)
      return fail(LYMPH_ERR_ARG, "tet_vert out of range");

  // polyhedron -> tets (ascending), block pattern in first-appearance order
  std::vector<int> ptr(np + 1, 0), ptet(nt);
  for (int t = 0; t < nt; ++t) ptr[tpoly[t] + 1]++;
  for (int i = 0; i < np; ++i) ptr[i + 1] += ptr[i];
  for (int i = 0; i < np; ++i)
    if (ptr[i + 1] == ptr[i]) return fail(LYMPH_EXIT_SETUP, "empty polyhedron " + std::to_string(i));
  {
    std::vector<int> fill(np, 0);
    for (int t = 0; t < nt; ++t) ptet[ptr[tpoly[t]] + fill[tpoly[t]]++] = t;
  }
  std::vector<int> row_ptr(np + 1), col, face_blk(4 * (size_t)nt, -1);
  col.reserve((size_t)np * 8);
  for (int P = 0; P < np; ++P) {
    row_ptr[P] = (int)col.size();
    col.push_back(P);
    for (int a = ptr[P]; a < ptr[P + 1]; ++a) {
      const int t = ptet[a];
      for (int f = 0; f < 4; ++f) {
        const int e2 = mesh->neigh[t * 4 + f];
        if (e2 < 0) continue;
        const int Q = tpoly[e2];
        if (Q == P) continue;
        int found = -1;
        for (int b = row_ptr[P] + 1; b < (int)col.size(); ++b)
          if (col[b] == Q) { found = b; break; }
        if (found < 0) {
          found = (int)col.size();
          col.push_back(Q);
        }
        face_blk[(size_t)t * 4 + f] = found;
      }
    }
  }
  row_ptr[np] = (int)col.size();

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(LYMPH_ERR_HIP, "no HIP device: liblymph_gpu needs an MI355X (gfx950)");

  LymphCtx* c = new LymphCtx();
  c->p = p;
  c->Np = np_of(p);
  c->n3 = 3 * c->Np;
  c->n3c = (c->n3 + 1) & ~1;
  c->nq3 = (p + 1) * (p + 1) * (p + 1);
  c->nq2 = (p + 1) * (p + 1);
  c->alpha = alpha;
  c->theta = theta;
  c->n_tet = nt;
  c->n_poly = np;
  c->n_blk = (int)col.size();
  c->h_row_ptr = row_ptr;
  c->h_tet_poly = tpoly;
  c->h_col = col;

  std::vector<double> txyz((size_t)nt * 12);
  for (int t = 0; t < nt; ++t)
    for (int v = 0; v < 4; ++v)
      for (int d = 0; d < 3; ++d)
        txyz[(size_t)t * 12 + v * 3 + d] = mesh->xyz[(size_t)mesh->tet_vert[t * 4 + v] * 3 + d];
  std::vector<double> qtet(c->nq3 * 4), wtet(c->nq3), qtri(c->nq2 * 4), wtri(c->nq2);
  quad_tables(p, qtet.data(), wtet.data(), qtri.data(), wtri.data());
  std::vector<int> bl(c->Np * 3);
  basis_list(p, bl.data());

  int rc = 0;
  auto up = [&](auto** dst, const auto* src, size_t n) {
    if (rc) return;
    rc = upload(dst, src, n);
    if (!rc) {
      c->owned.push_back((void*)*dst);
      c->bytes += n * sizeof(**dst);
    }
  };
  auto al = [&](auto** dst, size_t n) {
    if (rc) return;
    rc = dalloc(dst, n);
    if (!rc) {
      c->owned.push_back((void*)*dst);
      c->bytes += n * sizeof(**dst);
    }
  };
  MeshDev& m = c->md;
  double *d_txyz, *d_normal, *d_area, *d_rho, *d_lam, *d_mu, *d_bbox, *d_hk, *d_qtet, *d_wtet,
      *d_qtri, *d_wtri;
  int *d_tpoly, *d_neigh, *d_bc, *d_ptr, *d_ptet, *d_rowp, *d_col, *d_fblk, *d_bl;
  up(&d_txyz, txyz.data(), txyz.size());
  up(&d_tpoly, tpoly.data(), tpoly.size());
  up(&d_neigh, mesh->neigh, 4 * (size_t)nt);
  up(&d_bc, mesh->bc_space, 4 * (size_t)nt);
  up(&d_normal, mesh->normal, 12 * (size_t)nt);
  up(&d_area, mesh->area, 4 * (size_t)nt);
  up(&d_rho, mesh->rho, nt);
  up(&d_lam, mesh->lam, nt);
  up(&d_mu, mesh->mu, nt);
  up(&d_bbox, mesh->bbox, 6 * (size_t)np);
  up(&d_hk, mesh->hk, np);
  up(&d_ptr, ptr.data(), ptr.size());
  up(&d_ptet, ptet.data(), ptet.size());
  up(&d_rowp, row_ptr.data(), row_ptr.size());
  up(&d_col, col.data(), col.size());
  {
    std::vector<int> br(col.size());
    for (int P = 0; P < np; ++P) {
      c->max_row = std::max(c->max_row, row_ptr[P + 1] - row_ptr[P]);
      for (int b = row_ptr[P]; b < row_ptr[P + 1]; ++b) br[b] = P;
    }
    up(&c->blk_row, br.data(), br.size());
    // dispatch order for meshes of uneven polyhedra (longest row at least 1.5x
    // the mean): each XCD's contiguous share, longest block row first.  On
    // even meshes (cubes: rows of 4-7 blocks) the natural order is kept: there
    // the sort only moves boundary polyhedra, and measured slightly slower.
    if (2 * (size_t)c->max_row * np >= 3 * col.size()) {
      const int grid = (int)step_grid(np), per = grid / 8;
      std::vector<int> ord(grid, np);
      for (int x = 0; x < 8; ++x) {
        const int lo = x * per, hi = std::max(lo, std::min((x + 1) * per, np));
        for (int i = lo; i < hi; ++i) ord[i] = i;
        std::stable_sort(ord.begin() + lo, ord.begin() + hi, [&](int a, int b) {
          return row_ptr[a + 1] - row_ptr[a] > row_ptr[b + 1] - row_ptr[b];
        });
      }
      up(&c->order, ord.data(), ord.size());
    }
  }
  up(&d_fblk, face_blk.data(), face_blk.size());
  up(&d_qtet, qtet.data(), qtet.size());
  up(&d_wtet, wtet.data(), wtet.size());
  up(&d_qtri, qtri.data(), qtri.size());
  up(&d_wtri, wtri.data(), wtri.size());
  up(&d_bl, bl.data(), bl.size());
  const size_t npk = (size_t)c->Np * (c->Np + 1) / 2;
  al(&c->M, (size_t)np * c->Np * c->Np);
  al(&c->Mm, (size_t)np * c->Np * c->Np);
  al(&c->RM, (size_t)np * npk);
  al(&c->RMm, (size_t)np * npk);
  for (int s = 0; s < LYMPH_NUM_SLOTS; ++s) al(&c->slot[s], vec_n(c));
  for (int s = 0; s < 4; ++s) al(&c->stage[s], vec_n(c));
  al(&c->part, 3 * (size_t)np);
  al(&c->dsum, 2);
  al(&c->status, 1);
  if (rc) {
    lymph_ctx_destroy(c);
    return rc;
  }
  m.tet_xyz = d_txyz; m.tet_poly = d_tpoly; m.neigh = d_neigh; m.bc_space = d_bc;
  m.normal = d_normal; m.area = d_area; m.rho = d_rho; m.lam = d_lam; m.mu = d_mu;
  m.bbox = d_bbox; m.hk = d_hk; m.ptet_ptr = d_ptr; m.ptet = d_ptet; m.row_ptr = d_rowp;
  m.col = d_col; m.face_blk = d_fblk; m.qtet = d_qtet; m.wtet = d_wtet; m.qtri = d_qtri;
  m.wtri = d_wtri; m.blist = d_bl;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    lymph_ctx_destroy(c);
    return fail(LYMPH_ERR_HIP, "hipStreamCreate failed");
  }
  c->own_stream = true;
  for (int s = 0; s < LYMPH_NUM_SLOTS; ++s)
    if (hipMemsetAsync(c->slot[s], 0, vec_n(c) * sizeof(double), c->stream) != hipSuccess) {
      lymph_ctx_destroy(c);
      return fail(LYMPH_ERR_HIP, "hipMemsetAsync failed");
    }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    lymph_ctx_destroy(c);
    return fail(LYMPH_ERR_HIP, "stream sync failed");
  }
  // set-up of the allocation-free entry points: the scratch array's address
  // on this device, resolved once per device with the device current
  unsigned char* sp = nullptr;
  if (hipGetDevice(&c->device) != hipSuccess) {
    lymph_ctx_destroy(c);
    return fail(LYMPH_ERR_HIP, "hipGetDevice failed");
  }
  if (int r = scratch_ptr(c->device, &sp)) {
    lymph_ctx_destroy(c);
    return r;
  }
  *out = c;
  return 0;
}

int lymph_ctx_info(const LymphCtx* c, int64_t* info) {
  if (int rc = need_ctx(c)) return rc;
  const int64_t Np = c->Np, n3 = c->n3;
  info[0] = c->n_tet;
  info[1] = c->n_poly;
  info[2] = Np;
  info[3] = c->n_blk;
  info[4] = (int64_t)c->n_blk * (int64_t)blk_sz(c) * 8;
  info[5] = (int64_t)c->bytes;
  // B = 8 [ n_blk (3Np)^2 + N_poly Np(Np+1)/2 + 5 N_poly 3Np ]
  info[6] = 8 * ((int64_t)c->n_blk * n3 * n3 + (int64_t)c->n_poly * Np * (Np + 1) / 2 +
                 5 * (int64_t)c->n_poly * n3);
  // flops: 2 per operator entry + 2 triangular solves per component
  info[7] = 2 * (int64_t)c->n_blk * n3 * n3 + (int64_t)c->n_poly * 3 * 2 * Np * Np +
            (int64_t)c->n_poly * n3 * 6;
  return 0;
}

int lymph_ctx_set_layout(LymphCtx* c, int layout) {
  if (int rc = need_ctx(c)) return rc;
  if (layout != LYMPH_LAYOUT_POLY && layout != LYMPH_LAYOUT_FORTRAN)
    return fail(LYMPH_ERR_ARG, "unknown layout");
  c->layout = layout;
  return 0;
}

int lymph_ctx_set_stream(LymphCtx* c, void* stream) {
  if (int rc = need_ctx(c)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  c->own_stream = false;
  c->stream = (hipStream_t)stream;
  return 0;
}

void* lymph_ctx_get_stream(const LymphCtx* c) { return c ? (void*)c->stream : nullptr; }

int lymph_get_pattern(const LymphCtx* c, int32_t* row_ptr, int32_t* col) {
  if (int rc = need_ctx(c)) return rc;
  if (row_ptr) std::memcpy(row_ptr, c->h_row_ptr.data(), sizeof(int) * (c->n_poly + 1));
  if (col) std::memcpy(col, c->h_col.data(), sizeof(int) * c->n_blk);
  return 0;
}

int lymph_build_operator(LymphCtx* c) {
  if (int rc = need_ctx(c)) return rc;
  if (int rc = sym_off(c)) return rc;  // the blocks are about to change
  const size_t bs = blk_sz(c), nk = (size_t)c->n_blk * bs;
  // K (and A_dg unless dropped) live outside the mesh allocations so that
  // they can be released for batched runs
  if (!c->K) {
    if (int rc = dalloc(&c->K, nk)) return rc;
    c->bytes += nk * sizeof(double);
  }
  if (c->keep_adg && !c->A) {
    if (int rc = dalloc(&c->A, nk)) return rc;
    c->bytes += nk * sizeof(double);
  }
  c->built = false;
  HIP_TRY(hipMemsetAsync(c->K, 0, nk * sizeof(double), c->stream));
  if (c->A) HIP_TRY(hipMemsetAsync(c->A, 0, nk * sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(c->M, 0, (size_t)c->n_poly * c->Np * c->Np * sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(c->Mm, 0, (size_t)c->n_poly * c->Np * c->Np * sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const int chunk = 1 << 16;
    for (int p0 = 0; p0 < c->n_poly; p0 += chunk) {
      const int n = std::min(chunk, c->n_poly - p0);
      if (c->build_mode == LYMPH_BUILD_FACTORED)
        k_build_factored<P><<<n, factored_threads<P>(), 0, c->stream>>>(c->md, c->alpha, c->theta, c->K, c->A, c->M,
                                                      c->Mm, p0);
      else
        k_build_operator<P><<<n, 256, 0, c->stream>>>(c->md, c->alpha, c->theta, c->K, c->A, c->M,
                                                     c->Mm, p0);
      if (int r = check_launch("k_build_operator")) return r;
    }
    k_cholesky<P><<<(unsigned)c->n_poly, 64, 0, c->stream>>>(c->M, c->Mm, c->RM, c->RMm,
                                                              c->n_poly, c->status);
    if (int r = check_launch("k_cholesky")) return r;
    return 0;
  });
  if (rc) return rc;
  int st = 0;
  HIP_TRY(hipMemcpyAsync(&st, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (st) return fail(LYMPH_EXIT_SINGULARMTX, "mass matrix is not positive definite");
  if (c->K32)  // keep an existing 32-bit operator in step with the new blocks
    if (int r = build_ortho(c)) return r;
  c->built = true;
  return 0;
}

int lymph_ctx_set_build(LymphCtx* c, int mode) {
  if (int rc = need_ctx(c)) return rc;
  if (mode != LYMPH_BUILD_REFERENCE && mode != LYMPH_BUILD_FACTORED)
    return fail(LYMPH_ERR_ARG, "unknown build mode");
  c->build_mode = mode;
  return 0;
}

int lymph_ctx_set_keep_adg(LymphCtx* c, int keep) {
  if (int rc = need_ctx(c)) return rc;
  c->keep_adg = keep != 0;
  if (!c->keep_adg && c->A) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(c->A));
    c->A = nullptr;
    c->bytes -= (size_t)c->n_blk * blk_sz(c) * sizeof(double);
  }
  return 0;
}

int lymph_release_fp64_operator(LymphCtx* c) {
  if (int rc = need_built(c)) return rc;
  if (c->apply_bits == 64 || !c->K32)
    return fail(LYMPH_ERR_STATE, "select the 32- or 48-bit apply (lymph_set_apply_precision) first");
  if (c->K) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graph(c);
    HIP_TRY(hipFree(c->K));
    c->K = nullptr;
    c->bytes -= (size_t)c->n_blk * blk_sz(c) * sizeof(double);
  }
  return 0;
}

int lymph_set_symmetric_storage(LymphCtx* c, int on, double* asym) {
  if (int rc = need_built(c)) return rc;
  if (asym) *asym = 0.0;
  if (on != LYMPH_SYM_OFF && on != LYMPH_SYM_FP64 && on != LYMPH_SYM_FP64_FP32)
    return fail(LYMPH_ERR_ARG, "unknown storage mode");
  if (on == LYMPH_SYM_OFF) {
    c->sym32_wanted = false;
    return sym_off(c);
  }
  if (c->Ks && (on == LYMPH_SYM_FP64_FP32) == c->sym32_wanted) return 0;
  if (!c->K) return fail(LYMPH_ERR_STATE, "the FP64 operator was released");
  if (c->p > 3) return fail(LYMPH_ERR_ARG, "symmetric storage needs p <= 3");
  if (int rc = sym_off(c)) return rc;  // switching modes: start from the full rows
  c->sym32_wanted = on == LYMPH_SYM_FP64_FP32;
  const int rc = sym_on(c, kSymTol, asym);
  if (rc) c->sym32_wanted = false;
  return rc;
}

int lymph_export_operator(const LymphCtx* cc, double* K, double* A, double* M, double* Mm) {
  LymphCtx* c = const_cast<LymphCtx*>(cc);
  if (int rc = need_built(c)) return rc;
  const size_t bs = blk_sz(c), n3 = c->n3;
  std::vector<double> tmp((size_t)c->n_blk * bs);
  if (K && !c->K) return fail(LYMPH_ERR_STATE, "FP64 K blocks were released");
  if (A && !c->A) return fail(LYMPH_ERR_STATE, "A_dg blocks were not kept");
  for (int which = 0; which < 2; ++which) {
    double* dst = which ? A : K;
    if (!dst) continue;
    HIP_TRY(hipMemcpyAsync(tmp.data(), which ? c->A : c->K, tmp.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t b = 0; b < (size_t)c->n_blk; ++b)
      for (size_t r = 0; r < n3; ++r)
        for (size_t cl = 0; cl < n3; ++cl)
          dst[(b * n3 + r) * n3 + cl] = tmp[b * bs + kpos((int)n3, (int)r, (int)cl)];
  }
  const size_t mb = (size_t)c->n_poly * c->Np * c->Np * sizeof(double);
  if (M) HIP_TRY(hipMemcpyAsync(M, c->M, mb, hipMemcpyDeviceToHost, c->stream));
  if (Mm) HIP_TRY(hipMemcpyAsync(Mm, c->Mm, mb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_build_rhs(LymphCtx* c, int problem, int slot, double* rhs_out) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS) return fail(LYMPH_ERR_ARG, "bad slot");
  if (problem < LYMPH_PROB_NULL || problem > LYMPH_PROB_RICKER)
    return fail(LYMPH_ERR_ARG, "unknown problem");
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_rhs<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, problem, c->alpha, c->theta, c->slot[slot],
                                                nullptr, nullptr);
    return check_launch("k_rhs");
  });
  if (rc) return rc;
  if (rhs_out) return d2h_vec(c, c->slot[slot], rhs_out);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_project_fn(LymphCtx* c, lymph_field_fn fld, void* user, int slot, double* out) {
  if (int rc = need_built(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS) return fail(LYMPH_ERR_ARG, "bad slot");
  if (!fld) return fail(LYMPH_ERR_ARG, "field function is null");
  const int nt = c->n_tet, nq3 = c->nq3, nq2 = c->nq2;
  const size_t nv = (size_t)nt * nq3, nf = (size_t)nt * 4 * nq2;
  HIP_TRY(hipStreamSynchronize(c->stream));
  double *dv = nullptr, *df = nullptr;
  if (int rc = dalloc(&dv, 3 * nv)) return rc;
  if (int rc = dalloc(&df, 3 * nf)) {
    (void)hipFree(dv);
    return rc;
  }
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dv);
    (void)hipFree(df);
    return rc;
  };
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const size_t n = (size_t)nt * (nq3 + 4 * nq2);
    k_rhs_points<P><<<grid1(n, 256), 256, 0, c->stream>>>(c->md, nt, dv, df);
    return check_launch("k_rhs_points");
  });
  if (rc) return done(rc);
  HIP_TRY_OR(hipStreamSynchronize(c->stream), done);
  std::vector<double> vp;
  if ((rc = d2h(vp, dv, 3 * nv))) return done(rc);
  std::vector<double> tv(3 * nv, 0.0);
  fld((int)nv, vp.data(), tv.data(), user);
  HIP_TRY_OR(hipMemcpy(dv, tv.data(), 3 * nv * sizeof(double), hipMemcpyHostToDevice), done);
  rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_project<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, LYMPH_FN_ZERO, c->RMm, c->slot[slot], dv);
    return check_launch("k_project");
  });
  if (rc) return done(rc);
  if (out) rc = d2h_vec(c, c->slot[slot], out);
  return done(rc);
}

int lymph_build_rhs_fn(LymphCtx* c, lymph_forcing_fn f, lymph_dirichlet_fn gd, lymph_neumann_fn gn,
                       void* user, int slot, double* rhs_out) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS) return fail(LYMPH_ERR_ARG, "bad slot");
  const int nt = c->n_tet, nq3 = c->nq3, nq2 = c->nq2;
  const size_t nv = (size_t)nt * nq3, nf = (size_t)nt * 4 * nq2;
  HIP_TRY(hipStreamSynchronize(c->stream));
  double *dv = nullptr, *df = nullptr;
  if (int rc = dalloc(&dv, 3 * nv)) return rc;
  if (int rc = dalloc(&df, 3 * nf)) {
    (void)hipFree(dv);
    return rc;
  }
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dv);
    (void)hipFree(df);
    return rc;
  };
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const size_t n = (size_t)nt * (nq3 + 4 * nq2);
    k_rhs_points<P><<<grid1(n, 256), 256, 0, c->stream>>>(c->md, nt, dv, df);
    return check_launch("k_rhs_points");
  });
  if (rc) return done(rc);
  HIP_TRY_OR(hipStreamSynchronize(c->stream), done);
  std::vector<double> vp, fp, lam, mu, rho, nrm;
  std::vector<int> neigh, tag;
  if ((rc = d2h(vp, dv, 3 * nv)) || (rc = d2h(fp, df, 3 * nf)) || (rc = d2h(lam, c->md.lam, nt)) ||
      (rc = d2h(mu, c->md.mu, nt)) || (rc = d2h(rho, c->md.rho, nt)) ||
      (rc = d2h(nrm, c->md.normal, 12 * (size_t)nt)) || (rc = d2h(neigh, c->md.neigh, 4 * (size_t)nt)) ||
      (rc = d2h(tag, c->md.bc_space, 4 * (size_t)nt)))
    return done(rc);
  // forcing at every volume point (material of the point's tet)
  std::vector<double> tv(3 * nv, 0.0), tf(3 * nf, 0.0);
  if (f) {
    std::vector<double> l(nv), m(nv), r(nv);
    for (size_t i = 0; i < nv; ++i) {
      const size_t t = i / nq3;
      l[i] = lam[t];
      m[i] = mu[t];
      r[i] = rho[t];
    }
    f((int)nv, l.data(), m.data(), r.data(), vp.data(), tv.data(), user);
  }
  // Dirichlet (neighbour -1) and Neumann (-2) data at the boundary faces' points
  for (int kind : {-1, -2}) {
    if ((kind == -1 && !gd) || (kind == -2 && !gn)) continue;
    std::vector<size_t> at;  // index into the per-face point arrays
    for (size_t tf4 = 0; tf4 < 4 * (size_t)nt; ++tf4)
      if (neigh[tf4] == kind)
        for (int q = 0; q < nq2; ++q) at.push_back(tf4 * nq2 + q);
    if (at.empty()) continue;
    const size_t n = at.size();
    std::vector<double> x(3 * n), out(3 * n, 0.0), nn(3 * n), l(n), m(n);
    std::vector<int> tg(n);
    for (size_t k = 0; k < n; ++k) {
      const size_t tf4 = at[k] / nq2, t = tf4 / 4;
      for (int d = 0; d < 3; ++d) {
        x[3 * k + d] = fp[3 * at[k] + d];
        nn[3 * k + d] = nrm[3 * tf4 + d];
      }
      tg[k] = tag[tf4];
      l[k] = lam[t];
      m[k] = mu[t];
    }
    if (kind == -1)
      gd((int)n, x.data(), tg.data(), out.data(), user);
    else
      gn((int)n, l.data(), m.data(), nn.data(), x.data(), tg.data(), out.data(), user);
    for (size_t k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) tf[3 * at[k] + d] = out[3 * k + d];
  }
  HIP_TRY_OR(hipMemcpy(dv, tv.data(), 3 * nv * sizeof(double), hipMemcpyHostToDevice), done);
  HIP_TRY_OR(hipMemcpy(df, tf.data(), 3 * nf * sizeof(double), hipMemcpyHostToDevice), done);
  rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_rhs<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, LYMPH_PROB_NULL, c->alpha, c->theta,
                                                c->slot[slot], dv, df);
    return check_launch("k_rhs");
  });
  if (rc) return done(rc);
  if (rhs_out) rc = d2h_vec(c, c->slot[slot], rhs_out);
  return done(rc);
}

int lymph_project(LymphCtx* c, int fn, int slot, double* out) {
  if (int rc = need_built(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS) return fail(LYMPH_ERR_ARG, "bad slot");
  if (fn < LYMPH_FN_ZERO || fn > LYMPH_FN_CUBE_U) return fail(LYMPH_ERR_ARG, "unknown field");
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_project<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, fn, c->RMm, c->slot[slot], nullptr);
    return check_launch("k_project");
  });
  if (rc) return rc;
  if (out) return d2h_vec(c, c->slot[slot], out);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_set_slot(LymphCtx* c, int slot, const double* host) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS || !host) return fail(LYMPH_ERR_ARG, "bad slot/pointer");
  if (int rc = h2d_vec(c, host, c->slot[slot])) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_get_slot(LymphCtx* c, int slot, double* host) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS || !host) return fail(LYMPH_ERR_ARG, "bad slot/pointer");
  return d2h_vec(c, c->slot[slot], host);
}

double* lymph_slot_ptr(LymphCtx* c, int slot) {
  if (!c || slot < 0 || slot >= LYMPH_NUM_SLOTS) return nullptr;
  return c->slot[slot];
}

namespace {
// One step from device vectors in the original basis, time factor prepared:
// mode 1 (first step) a = u0, b = v0; mode 0 a = u^n, b = u^{n-1}.  The
// 32-bit apply steps the mass-orthonormal vectors and transforms back.
int step_dev(LymphCtx* c, int mode, double g, double c1, double c2, const double* a,
             const double* b, double* out, bool grouped) {
  if (c->apply_bits == 64) return launch_step(c, mode, g, c1, c2, a, b, out, grouped);
  if (int rc = tilde_rhs(c)) return rc;
  if (int rc = tilde(c, 0, a, c->tv[0])) return rc;
  if (int rc = tilde(c, 0, b, c->tv[1])) return rc;
  if (int rc = launch_step(c, mode, g, c1, c2, c->tv[0], c->tv[1], c->tv[2], grouped)) return rc;
  return tilde(c, 1, c->tv[2], out);
}

int sample_into(LymphCtx* c, const double* u, double* d_out);

// nsteps leapfrog steps on the slots U_PREV, U_CUR (t0 = time of the first
// of them); with `record`, receivers sampled every `every` steps.  The 64-bit
// apply rotates the slot pointers; the 32-bit apply steps mass-orthonormal
// copies and writes the last two states back into U_PREV, U_CUR.
int run_loop(LymphCtx* c, int problem, int nsteps, double t0, double dt, int every, bool record,
             double* t_out) {
  const bool b32 = c->apply_bits != 64;  // mass-orthonormal copies (32- and 48-bit apply)
  double *prev, *cur, *next;
  if (b32) {
    if (int rc = tilde_rhs(c)) return rc;
    if (int rc = tilde(c, 0, c->slot[LYMPH_SLOT_U_PREV], c->tv[0])) return rc;
    if (int rc = tilde(c, 0, c->slot[LYMPH_SLOT_U_CUR], c->tv[1])) return rc;
    prev = c->tv[0];
    cur = c->tv[1];
    next = c->tv[2];
  } else {
    prev = c->slot[LYMPH_SLOT_U_PREV];
    cur = c->slot[LYMPH_SLOT_U_CUR];
    next = c->slot[LYMPH_SLOT_U_NEXT];
  }
  double t = t0;
  const double dt2 = dt * dt;
  int rc = 0;
  for (int s = 1; s <= nsteps && !rc; ++s) {
    double g;
    bool grouped;
    rc = prepare_g(c, problem, t, &g, &grouped);
    if (!rc) rc = launch_step(c, 0, g, dt2, 0.0, cur, prev, next, grouped);
    double* tmp = prev;
    prev = cur;
    cur = next;
    next = tmp;
    t = t + dt;
    if (!rc && record && s % every == 0) {
      if (c->n_rec >= c->max_rec) {
        rc = fail(LYMPH_ERR_ARG, "receiver trace buffer full");
        break;
      }
      const double* src = cur;
      if (b32) {  // back to the original basis, receiver polyhedra only
        rc = tilde(c, 1, cur, c->stage[3], c->rpoly_u, c->n_rpoly_u);
        src = c->stage[3];
      }
      if (!rc) rc = sample_into(c, src, c->trace + (size_t)c->n_rec * 3 * c->n_recv);
      if (!rc) c->n_rec++;
    }
  }
  if (b32) {
    if (!rc) rc = tilde(c, 1, prev, c->slot[LYMPH_SLOT_U_PREV]);
    if (!rc) rc = tilde(c, 1, cur, c->slot[LYMPH_SLOT_U_CUR]);
  } else {
    c->slot[LYMPH_SLOT_U_PREV] = prev;
    c->slot[LYMPH_SLOT_U_CUR] = cur;
    c->slot[LYMPH_SLOT_U_NEXT] = next;
  }
  if (t_out) *t_out = t;
  return rc;
}
}  // namespace

int lymph_first_step_dev(LymphCtx* c, double g0, double dt, const double* u0, const double* v0,
                         double* u1) {
  if (int rc = need_built(c)) return rc;
  if (!u0 || !v0 || !u1) return fail(LYMPH_ERR_ARG, "null vector");
  const double dt2 = dt * dt, half_dt2 = 0.5 * dt2;
  return step_dev(c, 1, g0, half_dt2, dt, u0, v0, u1, false);
}

int lymph_step_dev(LymphCtx* c, double g, double dt, const double* um1, const double* u,
                   double* up1) {
  if (int rc = need_built(c)) return rc;
  if (!um1 || !u || !up1) return fail(LYMPH_ERR_ARG, "null vector");
  return step_dev(c, 0, g, dt * dt, 0.0, u, um1, up1, false);
}

int lymph_first_step(LymphCtx* c, double g0, double dt, const double* u0, const double* v0,
                     double* u1) {
  if (int rc = need_built(c)) return rc;
  if (!u0 || !v0 || !u1) return fail(LYMPH_ERR_ARG, "null vector");
  if (int rc = h2d_vec(c, u0, c->stage[0])) return rc;
  if (int rc = h2d_vec(c, v0, c->stage[1])) return rc;
  if (int rc = lymph_first_step_dev(c, g0, dt, c->stage[0], c->stage[1], c->stage[2])) return rc;
  return d2h_vec(c, c->stage[2], u1);
}

int lymph_step(LymphCtx* c, double g, double dt, const double* um1, const double* u, double* up1) {
  if (int rc = need_built(c)) return rc;
  if (!um1 || !u || !up1) return fail(LYMPH_ERR_ARG, "null vector");
  if (int rc = h2d_vec(c, um1, c->stage[0])) return rc;
  if (int rc = h2d_vec(c, u, c->stage[1])) return rc;
  if (int rc = lymph_step_dev(c, g, dt, c->stage[0], c->stage[1], c->stage[2])) return rc;
  return d2h_vec(c, c->stage[2], up1);
}

int lymph_step_n(LymphCtx* c, int problem, int nsteps, double t0, double dt, double* t_out) {
  if (int rc = need_built(c)) return rc;
  if (nsteps < 0) return fail(LYMPH_ERR_ARG, "nsteps < 0");
  return run_loop(c, problem, nsteps, t0, dt, 1, false, t_out);
}

int lymph_time_steps(LymphCtx* c, int problem, int nsteps, double t0, double dt, double* ms) {
  if (int rc = need_built(c)) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float f = 0.f;
  int rc = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipEventRecord(e0, c->stream) != hipSuccess) {
    rc = fail(LYMPH_ERR_HIP, "event set-up failed");
  } else {
    rc = lymph_step_n(c, problem, nsteps, t0, dt, nullptr);
    if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&f, e0, e1) != hipSuccess)
      if (!rc) rc = fail(LYMPH_ERR_HIP, "event timing failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (ms) *ms = f;
  return 0;
}

int lymph_autotune(LymphCtx* c, int reps, int* n_variants, int* unr, int* xmode, int* wpp,
                   double* ms, int* chosen) {
  if (int rc = need_built(c)) return rc;
  if (reps < 1) return fail(LYMPH_ERR_ARG, "reps < 1");
  // with face pairs stored once the step runs k_sym_* whatever the variant:
  // timing the variants would only time those kernels again
  if ((c->apply_bits == 64 && c->Ks) || (c->apply_bits == 32 && c->Ks32) ||
      (c->apply_bits == 48 && c->Ks16))
    return fail(LYMPH_ERR_STATE, "face pairs are stored once: the full-row variants are not in use");
  return dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const int bits = c->apply_bits;
    const int NV = n_var_of<P>(bits);
    if (n_variants) *n_variants = NV;
    const bool fits = (size_t)c->max_row * c->n3 * sizeof(double) <= kMaxDynLds;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // scratch inputs: the current state slots are copied so nothing the
    // caller holds is modified
    const size_t nb = vec_n(c) * sizeof(double);
    HIP_TRY(hipMemcpyAsync(c->stage[0], c->slot[LYMPH_SLOT_U_PREV], nb, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->stage[1], c->slot[LYMPH_SLOT_U_CUR], nb, hipMemcpyDeviceToDevice, c->stream));
    const int saved = c->step_variant;
    double best = 1e300;
    int bi = 0;
    for (int v = 0; v < NV; ++v) {
      const int w = bits == 64 ? StepVar<P>::wpp[v]
                               : (bits == 32 ? StepVar32<P>::wpp[v] : StepVar48<P>::wpp[v]);
      if (unr)
        unr[v] = bits == 64 ? StepVar<P>::unr[v]
                            : (bits == 32 ? StepVar32<P>::unr[v] : StepVar48<P>::unr[v]);
      if (xmode)
        xmode[v] = bits == 64 ? StepVar<P>::xs[v]
                              : 1 + (bits == 32 ? StepVar32<P>::pk[v] : StepVar48<P>::pk[v]);
      if (wpp) wpp[v] = w;
      if (w > 1 && !fits) {  // would run variant 0: not a candidate
        if (ms) ms[v] = -1.0;
        continue;
      }
      c->step_variant = v;
      if (int rc = launch_step(c, 0, 0.5, 1e-8, 0.0, c->stage[1], c->stage[0], c->stage[2])) {
        c->step_variant = saved;
        return rc;
      }
      HIP_TRY(hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r)
        if (int rc = launch_step(c, 0, 0.5, 1e-8, 0.0, c->stage[1], c->stage[0], c->stage[2])) {
          c->step_variant = saved;
          return rc;
        }
      HIP_TRY(hipEventRecord(e1, c->stream));
      HIP_TRY(hipEventSynchronize(e1));
      float f = 0.f;
      HIP_TRY(hipEventElapsedTime(&f, e0, e1));
      const double t = (double)f / reps;
      if (ms) ms[v] = t;
      if (t < best) { best = t; bi = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->step_variant = bi;
    if (chosen) *chosen = bi;
    return 0;
  });
}

int lymph_set_step_variant(LymphCtx* c, int index, int* n_variants) {
  if (int rc = need_ctx(c)) return rc;
  return dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const int NV = n_var_of<P>(c->apply_bits);
    if (n_variants) *n_variants = NV;
    if (index < 0 || index >= NV) return fail(LYMPH_ERR_ARG, "no such step variant");
    c->step_variant = index;
    return 0;
  });
}

int lymph_set_apply_precision(LymphCtx* c, int bits) {
  if (int rc = need_built(c)) return rc;
  if (bits != 32 && bits != 48 && bits != 64)
    return fail(LYMPH_ERR_ARG, "precision must be 32, 48 or 64");
  if (bits == 64 && !c->K)
    return fail(LYMPH_ERR_STATE, "FP64 K blocks were released: rebuild the operator");
  if (bits == 48 && !c->K16 && !c->K)
    return fail(LYMPH_ERR_STATE, "the 48-bit corrections need the FP64 K blocks: rebuild the operator");
  if (bits != c->apply_bits) drop_graph(c);  // a captured solve steps the other apply
  const size_t n16 = (size_t)c->n_blk * Dims16Any(c->p);
  auto drop16 = [&]() {  // corrections allocated by this call but never written
    (void)hipFree(c->K16);
    c->K16 = nullptr;
    c->bytes -= n16 * sizeof(_Float16);
  };
  bool new16 = false;
  if (bits == 48 && !c->K16) {
    // the corrections come with a fresh ~K: both are written by one pass
    _Float16* k16 = nullptr;
    if (int r = dalloc(&k16, n16)) return r;
    c->K16 = k16;
    c->bytes += n16 * sizeof(_Float16);
    new16 = true;
    if (c->K32)
      if (int r = build_ortho(c)) {
        drop16();
        return r;
      }
  }
  if (bits != 64 && !c->K32) {
    // vectors first: K32 exists only once everything the 32-bit path uses does
    for (double*& q : c->tv) {
      if (!q) {
        if (int r = dalloc(&q, vec_n(c))) {
          if (new16) drop16();
          return r;
        }
        c->bytes += vec_n(c) * sizeof(double);
      }
    }
    const size_t n = (size_t)c->n_blk * Dims32Any(c->p);
    float* k32 = nullptr;
    if (int r = dalloc(&k32, n)) {
      if (new16) drop16();
      return r;
    }
    c->K32 = k32;
    c->bytes += n * sizeof(float);
    if (int r = build_ortho(c)) {
      (void)hipFree(c->K32);
      c->K32 = nullptr;
      c->bytes -= n * sizeof(float);
      if (new16) drop16();
      return r;
    }
  }
  if (bits != 64)
    if (int rc = sym32_build(c)) return rc;  // pairs stored once: the 32- and 48-bit copies too
  c->apply_bits = bits;
  c->step_variant = 0;
  return 0;
}

int lymph_set_source_groups(LymphCtx* c, int n_groups, const int32_t* poly_group,
                            const double* f0, const double* t0) {
  if (int rc = need_ctx(c)) return rc;
  if (n_groups < 0) return fail(LYMPH_ERR_ARG, "n_groups < 0");
  if (n_groups > 0) {
    if (!poly_group || !f0 || !t0) return fail(LYMPH_ERR_ARG, "null group array");
    for (int i = 0; i < c->n_poly; ++i)
      if (poly_group[i] < 0 || poly_group[i] >= n_groups)
        return fail(LYMPH_ERR_ARG, "polyhedron group out of range");
    for (int i = 0; i < n_groups; ++i)
      if (!(f0[i] > 0)) return fail(LYMPH_ERR_ARG, "f0 must be positive");
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  drop_graph(c);  // a captured solve reads these buffers by address
  for (void** q : {(void**)&c->gsel, (void**)&c->gf0, (void**)&c->gt0, (void**)&c->gtab})
    if (*q) {
      (void)hipFree(*q);
      *q = nullptr;
    }
  c->n_groups = 0;
  if (n_groups == 0) return 0;
  if (int rc = upload(&c->gsel, poly_group, (size_t)c->n_poly)) return rc;
  if (int rc = upload(&c->gf0, f0, (size_t)n_groups)) return rc;
  if (int rc = upload(&c->gt0, t0, (size_t)n_groups)) return rc;
  if (int rc = dalloc(&c->gtab, (size_t)n_groups)) return rc;
  c->n_groups = n_groups;
  return 0;
}

int lymph_set_ricker(LymphCtx* c, double f0, double t0) {
  if (int rc = need_ctx(c)) return rc;
  if (!(f0 > 0)) return fail(LYMPH_ERR_ARG, "f0 must be positive");
  if (f0 != c->ricker_f0 || t0 != c->ricker_t0) {
    // a captured solve holds the old wavelet's factors as kernel arguments
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graph(c);
  }
  c->ricker_f0 = f0;
  c->ricker_t0 = t0;
  return 0;
}

double lymph_ricker(double f0, double t0, double t) { return ricker(f0, t0, t); }

int lymph_add_point_sources(LymphCtx* c, int n, const double* xyz, const int32_t* tet,
                            const double* force, int slot) {
  if (int rc = need_ctx(c)) return rc;
  if (n < 0 || slot < 0 || slot >= LYMPH_NUM_SLOTS || (n > 0 && (!xyz || !tet || !force)))
    return fail(LYMPH_ERR_ARG, "bad point-source arguments");
  if (n == 0) return 0;
  std::vector<int> sp(n);
  for (int i = 0; i < n; ++i) {
    if (tet[i] < 0 || tet[i] >= c->n_tet) return fail(LYMPH_ERR_ARG, "source tet out of range");
    sp[i] = c->h_tet_poly[tet[i]];
  }
  double *dx = nullptr, *df = nullptr;
  int* dp = nullptr;
  int rc = upload(&dx, xyz, 3 * (size_t)n);
  if (!rc) rc = upload(&df, force, 3 * (size_t)n);
  if (!rc) rc = upload(&dp, sp.data(), sp.size());
  if (!rc)
    rc = dispatch_p(c->p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      k_point_sources<P><<<1, 128, 0, c->stream>>>(dx, dp, df, n, c->md.bbox, c->md.blist,
                                                  c->slot[slot], nullptr);
      return check_launch("k_point_sources");
    });
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(LYMPH_ERR_HIP, "sync failed");
  if (dx) (void)hipFree(dx);
  if (df) (void)hipFree(df);
  if (dp) (void)hipFree(dp);
  return rc;
}

int lymph_add_double_couples(LymphCtx* c, int n, const int32_t* tet, const double* strike,
                             const double* normal, const double* m0, int slot) {
  if (int rc = need_ctx(c)) return rc;
  if (n < 0 || slot < 0 || slot >= LYMPH_NUM_SLOTS || (n > 0 && (!tet || !strike || !normal || !m0)))
    return fail(LYMPH_ERR_ARG, "bad double-couple arguments");
  if (n == 0) return 0;
  for (int i = 0; i < n; ++i)
    if (tet[i] < 0 || tet[i] >= c->n_tet) return fail(LYMPH_ERR_ARG, "source tet out of range");
  int* dt = nullptr;
  double *ds = nullptr, *dn = nullptr, *dm = nullptr;
  int rc = upload(&dt, tet, (size_t)n);
  if (!rc) rc = upload(&ds, strike, 3 * (size_t)n);
  if (!rc) rc = upload(&dn, normal, 3 * (size_t)n);
  if (!rc) rc = upload(&dm, m0, (size_t)n);
  if (!rc)
    rc = dispatch_p(c->p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      k_double_couples<P><<<1, 128, 0, c->stream>>>(c->md, dt, ds, dn, dm, n, c->slot[slot]);
      return check_launch("k_double_couples");
    });
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(LYMPH_ERR_HIP, "sync failed");
  for (void* q : {(void*)dt, (void*)ds, (void*)dn, (void*)dm})
    if (q) (void)hipFree(q);
  return rc;
}

int lymph_add_point_sources_noalloc(LymphCtx* c, int n, const double* xyz, const int32_t* tet,
                                    const double* force, int slot) {
  if (int rc = need_ctx(c)) return rc;
  if (n < 0 || slot < 0 || slot >= LYMPH_NUM_SLOTS || (n > 0 && (!xyz || !tet || !force)))
    return fail(LYMPH_ERR_ARG, "bad point-source arguments");
  if (n > kNoAllocMaxSources)
    return fail(LYMPH_ERR_ARG, "lymph_add_point_sources_noalloc takes at most " +
                                   std::to_string(kNoAllocMaxSources) + " sources per call, got " +
                                   std::to_string(n));
  if (n == 0) return 0;
  for (int i = 0; i < n; ++i)
    if (tet[i] < 0 || tet[i] >= c->n_tet) return fail(LYMPH_ERR_ARG, "source tet out of range");
  int cur = -1;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != c->device) return fail(LYMPH_ERR_STATE, "the context's device is not the current device");
  unsigned char* base = nullptr;
  if (int rc = scratch_ptr(c->device, &base)) return rc;
  double* sx = (double*)base;
  double* sf = sx + 3 * (size_t)kNoAllocMaxSources;
  int* st = (int*)(sf + 3 * (size_t)kNoAllocMaxSources);
  HIP_TRY(hipMemcpyAsync(sx, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(sf, force, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(st, tet, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_point_sources<P><<<1, 128, 0, c->stream>>>(sx, st, sf, n, c->md.bbox, c->md.blist,
                                                c->slot[slot], c->md.tet_poly);
    return check_launch("k_point_sources");
  });
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host arrays are only borrowed
  return 0;
}

int lymph_set_receivers(LymphCtx* c, int n, const double* xyz, const int32_t* tet,
                        int max_records) {
  if (int rc = need_ctx(c)) return rc;
  if (n < 0 || max_records < 0 || (n > 0 && (!xyz || !tet)))
    return fail(LYMPH_ERR_ARG, "bad receiver arguments");
  std::vector<int> rp(n);
  for (int i = 0; i < n; ++i) {
    if (tet[i] < 0 || tet[i] >= c->n_tet) return fail(LYMPH_ERR_ARG, "receiver tet out of range");
    rp[i] = c->h_tet_poly[tet[i]];
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  drop_graph(c);  // a captured solve samples into these buffers by address
  if (c->rpoly) (void)hipFree(c->rpoly);
  if (c->rxyz) (void)hipFree(c->rxyz);
  if (c->rnow) (void)hipFree(c->rnow);
  if (c->trace) (void)hipFree(c->trace);
  if (c->rpoly_u) (void)hipFree(c->rpoly_u);
  c->rpoly = nullptr; c->rxyz = nullptr; c->rnow = nullptr; c->trace = nullptr;
  c->rpoly_u = nullptr; c->n_rpoly_u = 0;
  c->n_recv = 0; c->max_rec = 0; c->n_rec = 0;
  if (n == 0) return 0;
  // receivers count only once every buffer they use exists
  if (int rc = upload(&c->rpoly, rp.data(), rp.size())) return rc;
  if (int rc = upload(&c->rxyz, xyz, 3 * (size_t)n)) return rc;
  if (int rc = dalloc(&c->rnow, 3 * (size_t)n)) return rc;
  {
    std::vector<int> u(rp);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    if (int rc = upload(&c->rpoly_u, u.data(), u.size())) return rc;
    c->n_rpoly_u = (int)u.size();
  }
  if (int rc = dalloc(&c->trace, 3 * (size_t)n * (size_t)(max_records > 0 ? max_records : 1))) return rc;
  c->n_recv = n;
  c->max_rec = max_records;
  return 0;
}

namespace {
int sample_into(LymphCtx* c, const double* u, double* d_out) {
  return dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_sample<P><<<grid1(3 * (size_t)c->n_recv, 128), 128, 0, c->stream>>>(
        u, c->rpoly, c->rxyz, c->md.bbox, c->md.blist, c->n_recv, d_out);
    return check_launch("k_sample");
  });
}
}  // namespace

int lymph_sample_receivers(LymphCtx* c, int slot, double* out) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS || !out) return fail(LYMPH_ERR_ARG, "bad slot/pointer");
  if (c->n_recv == 0) return fail(LYMPH_ERR_STATE, "no receivers set");
  if (int rc = sample_into(c, c->slot[slot], c->rnow)) return rc;
  HIP_TRY(hipMemcpyAsync(out, c->rnow, 3 * (size_t)c->n_recv * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_step_n_record(LymphCtx* c, int problem, int nsteps, double t0, double dt, int every,
                        double* t_out) {
  if (int rc = need_built(c)) return rc;
  if (c->n_recv == 0) return fail(LYMPH_ERR_STATE, "no receivers set");
  if (every < 1 || nsteps < 0) return fail(LYMPH_ERR_ARG, "bad nsteps/every");
  return run_loop(c, problem, nsteps, t0, dt, every, true, t_out);
}

int lymph_get_traces(LymphCtx* c, double* out, int* n_records) {
  if (int rc = need_ctx(c)) return rc;
  if (n_records) *n_records = c->n_rec;
  if (out && c->n_rec > 0)
    HIP_TRY(hipMemcpyAsync(out, c->trace, (size_t)c->n_rec * 3 * c->n_recv * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_capture_solve(LymphCtx* c, int problem, double dt, int nsteps, int every) {
  if (int rc = need_built(c)) return rc;
  if (nsteps < 0 || every < 1) return fail(LYMPH_ERR_ARG, "bad nsteps/every");
  if (c->n_recv > 0 && nsteps / every > c->max_rec)
    return fail(LYMPH_ERR_ARG, "receiver trace buffer too small for the captured solve");
  if (!c->own_stream) return fail(LYMPH_ERR_STATE, "capture needs the context's own stream");
  drop_graph(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->n_rec = 0;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  // zero initial state (projection kernels, no memset nodes), first step at
  // t = 0, then the loop; every launch goes to the capturing stream
  int rc = dispatch_p(c->p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_project<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, LYMPH_FN_ZERO, c->RMm,
                                                   c->slot[LYMPH_SLOT_U_PREV], nullptr);
    k_project<P><<<c->n_poly, 128, 0, c->stream>>>(c->md, LYMPH_FN_ZERO, c->RMm,
                                                   c->slot[LYMPH_SLOT_V0], nullptr);
    return check_launch("k_project");
  });
  double t = 0.0;
  double g0 = 0.0;
  bool grouped = false;
  if (!rc) rc = prepare_g(c, problem, t, &g0, &grouped);
  if (!rc)
    rc = step_dev(c, 1, g0, 0.5 * (dt * dt), dt, c->slot[LYMPH_SLOT_U_PREV],
                  c->slot[LYMPH_SLOT_V0], c->slot[LYMPH_SLOT_U_CUR], grouped);
  t = t + dt;
  if (!rc) {
    if (c->n_recv > 0) rc = lymph_step_n_record(c, problem, nsteps, t, dt, every, &t);
    else rc = lymph_step_n(c, problem, nsteps, t, dt, &t);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(LYMPH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  c->graph = g;
  HIP_TRY(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  c->graph_records = c->n_rec;
  for (int k = 0; k < LYMPH_NUM_SLOTS; ++k) c->graph_slot[k] = c->slot[k];
  return 0;
}

int lymph_replay(LymphCtx* c) {
  if (int rc = need_ctx(c)) return rc;
  if (!c->graph_exec) return fail(LYMPH_ERR_STATE, "no captured solve: call lymph_capture_solve");
  HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
  c->n_rec = c->graph_records;
  // the 64-bit loop rotates the state slots; the replay ends where the
  // captured solve ended, whatever uncaptured steps rotated since
  for (int k = 0; k < LYMPH_NUM_SLOTS; ++k) c->slot[k] = c->graph_slot[k];
  return 0;
}

int lymph_sync(LymphCtx* c) {
  if (int rc = need_ctx(c)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int lymph_error_l2(LymphCtx* c, const double* uh, const double* uex, double* err2) {
  if (int rc = need_built(c)) return rc;
  if (int rc = h2d_vec(c, uh, c->stage[0])) return rc;
  if (int rc = h2d_vec(c, uex, c->stage[1])) return rc;
  return errors_dev(c, c->stage[0], c->stage[1], err2, nullptr);
}

int lymph_error_dg(LymphCtx* c, const double* uh, const double* uex, double* err2) {
  if (int rc = need_built(c)) return rc;
  if (int rc = h2d_vec(c, uh, c->stage[0])) return rc;
  if (int rc = h2d_vec(c, uex, c->stage[1])) return rc;
  return errors_dev(c, c->stage[0], c->stage[1], nullptr, err2);
}

int lymph_error_slots(LymphCtx* c, int su, int se, double* l2sq, double* dgsq) {
  if (int rc = need_built(c)) return rc;
  if (su < 0 || su >= LYMPH_NUM_SLOTS || se < 0 || se >= LYMPH_NUM_SLOTS)
    return fail(LYMPH_ERR_ARG, "bad slot");
  return errors_dev(c, c->slot[su], c->slot[se], l2sq, dgsq);
}

namespace {
// Instability control (the reference's EXIT_INSTAB): *finite = 0 if the
// vector holds a NaN or an infinity.  Synchronises the context's stream.
int finite_dev(LymphCtx* c, const double* v, int* finite) {
  const size_t n = vec_n(c);
  k_set_int<<<1, 1, 0, c->stream>>>(c->status, 0);
  k_nonfinite<<<(unsigned)std::min<size_t>(grid1(n, 256), 2048), 256, 0, c->stream>>>(v, n, c->status);
  if (int rc = check_launch("k_nonfinite")) return rc;
  int h = 0;
  HIP_TRY(hipMemcpyAsync(&h, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *finite = h ? 0 : 1;
  return 0;
}
constexpr int kInstabEvery = 256;  // lymph_run checks u^n every this many steps
}  // namespace

int lymph_check_finite(LymphCtx* c, int slot, int* finite) {
  if (int rc = need_ctx(c)) return rc;
  if (slot < 0 || slot >= LYMPH_NUM_SLOTS) return fail(LYMPH_ERR_ARG, "bad slot");
  if (!finite) return fail(LYMPH_ERR_ARG, "finite is null");
  return finite_dev(c, c->slot[slot], finite);
}

int lymph_run(LymphCtx* c, int problem, double dt, double stop_time, int max_steps,
              double* u_final, double* out) {
  if (int rc = need_ctx(c)) return rc;
  if (!out) return fail(LYMPH_ERR_ARG, "out is null");
  if (!c->built)
    if (int rc = lymph_build_operator(c)) return rc;
  if (int rc = lymph_build_rhs(c, LYMPH_PROB_NULL, LYMPH_SLOT_RHS_STAT, nullptr)) return rc;
  if (int rc = lymph_build_rhs(c, problem, LYMPH_SLOT_RHS_DYN, nullptr)) return rc;
  const int fn_u0 = problem == LYMPH_PROB_CUBE_WAVE ? LYMPH_FN_CUBE_U : LYMPH_FN_ZERO;
  const int fn_v0 = problem == LYMPH_PROB_REF_MMS ? LYMPH_FN_REF_VEL : LYMPH_FN_ZERO;
  const int fn_ex = problem == LYMPH_PROB_REF_MMS
                        ? LYMPH_FN_REF_UEX
                        : (problem == LYMPH_PROB_CUBE_WAVE ? LYMPH_FN_CUBE_U : LYMPH_FN_ZERO);
  if (int rc = lymph_project(c, fn_u0, LYMPH_SLOT_U_PREV, nullptr)) return rc;
  if (int rc = lymph_project(c, fn_v0, LYMPH_SLOT_V0, nullptr)) return rc;
  double t = 0.0;
  int nsteps = 0;
  if (int rc = lymph_first_step_dev(c, time_factor(problem, t), dt, c->slot[LYMPH_SLOT_U_PREV],
                                    c->slot[LYMPH_SLOT_V0], c->slot[LYMPH_SLOT_U_CUR]))
    return rc;
  nsteps++;
  t = t + dt;
  while (t <= stop_time && (max_steps < 0 || nsteps < max_steps)) {
    double tn;
    if (int rc = lymph_step_n(c, problem, 1, t, dt, &tn)) return rc;
    nsteps++;
    t = tn;
    if (nsteps % kInstabEvery == 0 || !(t <= stop_time) || nsteps == max_steps) {
      int ok = 1;
      if (int rc = finite_dev(c, c->slot[LYMPH_SLOT_U_CUR], &ok)) return rc;
      if (!ok)
        return fail(LYMPH_EXIT_INSTAB, "numerical instability: u holds NaN/Inf after step " +
                                           std::to_string(nsteps) + " (t = " + std::to_string(t) + ")");
    }
  }
  if (int rc = lymph_project(c, fn_ex, LYMPH_SLOT_UEX, nullptr)) return rc;
  k_scale<<<grid1(vec_n(c), 256), 256, 0, c->stream>>>(c->slot[LYMPH_SLOT_UEX],
                                                       time_factor(problem, t), vec_n(c));
  if (int rc = check_launch("k_scale")) return rc;
  double l2, dg;
  if (int rc = errors_dev(c, c->slot[LYMPH_SLOT_U_CUR], c->slot[LYMPH_SLOT_UEX], &l2, &dg)) return rc;
  if (u_final)
    if (int rc = d2h_vec(c, c->slot[LYMPH_SLOT_U_CUR], u_final)) return rc;
  out[0] = t;
  out[1] = nsteps;
  out[2] = std::sqrt(l2);
  out[3] = std::sqrt(dg);
  return 0;
}

}  // extern "C"
